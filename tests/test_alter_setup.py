"""&alter_setup: harmonic and subharmonic conversion of beam and fields
(src/Main/AlterSetup.cpp:37-220, Beam.cpp:187-299, Field.cpp:188-288).

The oracle's restatement of the four conversions is pinned here by what
the conversions must preserve: the window order of the field ring (window
slice w lives in storage slot (w + first) % n), the round trip of an up-
then down-conversion, the merged slice as the power-normalised sum of its
group, and every particle's absolute longitudinal position through a
resampling conversion and the following sort.  The deck runner's own
conversions are then checked against the oracle's on the CPU (check mode,
dumps before and after) and, on the GPU, through a tracked run after the
conversion.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "genesis-1.3-version4_amd", "python"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import g4mi  # noqa: E402
import oracle_py as O  # noqa: E402

TWO_PI = 4.0 * np.arcsin(1.0)


def _window(a, first):
    n = a.shape[0]
    return np.array([a[(w + first) % n] for w in range(n)])


def test_field_harmonic_conversion_window_order_and_round_trip():
    """Up-conversion copies window slice w into window slices h w .. h w + h - 1
    for every ring offset, and down-conversion by the same h restores the
    field and its offset."""
    rng = np.random.default_rng(3)
    ns, cells, h = 6, 5, 3
    for first in range(ns):
        f = rng.standard_normal((ns, cells, 2))
        up, fu = O.field_harmonic_conversion(f, first, h, True)
        assert up.shape[0] == ns * h
        assert np.array_equal(_window(up, fu), np.repeat(_window(f, first), h, axis=0))
        back, fb = O.field_subharmonic_conversion(up, fu, h, True)
        assert fb == first
        assert np.allclose(back, f, rtol=1e-14, atol=1e-14)
        same, fs = O.field_harmonic_conversion(f, first, h, False)
        assert fs == first and np.array_equal(same, f)


@pytest.mark.parametrize("h", [2, 3])
def test_field_subharmonic_merge_in_window_order(h):
    """Each new window slice is the sum of its h old window slices, scaled so
    that its power is the group's mean power."""
    rng = np.random.default_rng(11 + h)
    ns, cells = 6, 7
    for first in range(ns):
        g = rng.standard_normal((ns, cells, 2))
        sub, fs = O.field_subharmonic_conversion(g, first, h, True)
        W, S = _window(g, first), _window(sub, fs)
        assert S.shape[0] == ns // h
        for w in range(ns // h):
            grp = W[w * h:(w + 1) * h]
            s = grp.sum(axis=0)
            pm = np.mean([np.sum(x * x) for x in grp])
            assert np.allclose(S[w], s * np.sqrt(pm / np.sum(s * s)), rtol=1e-13, atol=1e-13)


def _one4one_beam(rng, counts, sample):
    off = np.zeros(len(counts) + 1, np.int64)
    off[1:] = np.cumsum(counts)
    n = int(off[-1])
    theta = rng.uniform(0.0, TWO_PI * sample, n)
    return off, theta, [rng.standard_normal(n) for _ in range(5)]


def _positions(off, theta, lam, sample):
    """Absolute longitudinal positions [m]: (a 2 pi sample + theta) lam / 2 pi."""
    a = np.repeat(np.arange(len(off) - 1), np.diff(off))
    return (a * TWO_PI * sample + theta) * lam / TWO_PI


def _sort(off, theta, others, lam, sample):
    g, x, y, px, py = (v.copy() for v in others)
    th = theta.copy()
    cur = np.zeros(len(off) - 1)
    off = off.copy()
    slen = TWO_PI * sample  # 2 pi slicelength / reflength
    kept = O.lib().orc_sort_one4one(len(off) - 1, off.ctypes.data_as(O._lp), len(th), slen, lam * sample,
                                    *[O._ptr(v) for v in (g, th, x, y, px, py, cur)])
    assert kept >= 0
    n = int(off[-1])
    return off, th[:n], g[:n], cur


@pytest.mark.parametrize("h,sample", [(2, 1), (3, 2)])
def test_beam_resampling_keeps_positions(h, sample):
    """Harmonic up-conversion with resampling and its sort, and subharmonic
    down-conversion with resampling and its sort, keep every particle and
    its absolute position (to rounding); the currents are n ce / slicelength
    of the new slices."""
    rng = np.random.default_rng(5 + h)
    lam = 1e-10
    off, theta, others = _one4one_beam(rng, [40, 0, 55, 31, 60, 12], sample)
    s0 = np.sort(_positions(off, theta, lam, sample))
    o1, t1 = O.beam_harmonic_conversion(off, theta, h, True)
    assert len(o1) - 1 == 6 * h
    o1, t1, g1, cur = _sort(o1, t1, others, lam / h, sample)
    assert o1[-1] == off[-1]
    assert np.allclose(np.sort(_positions(o1, t1, lam / h, sample)), s0, rtol=1e-14, atol=0)
    assert np.allclose(cur, np.diff(o1) * O.CE / (lam / h * sample), rtol=1e-15)
    # and back down to the original slices
    o2, t2 = O.beam_subharmonic_conversion(o1, t1, h, True, lam / h, sample)
    assert len(o2) - 1 == 6
    o2, t2, g2, cur2 = _sort(o2, t2, [g1] * 5, lam, sample)
    assert o2[-1] == off[-1]
    assert np.array_equal(np.diff(o2), np.diff(off))
    assert np.allclose(np.sort(_positions(o2, t2, lam, sample)), s0, rtol=1e-14, atol=0)


def test_beam_conversion_without_resampling_scales_theta():
    rng = np.random.default_rng(1)
    off, theta, _ = _one4one_beam(rng, [8, 9], 1)
    o, t = O.beam_harmonic_conversion(off, theta, 5, False)
    assert np.array_equal(o, off) and np.array_equal(t, theta * 5.0)
    o, t = O.beam_subharmonic_conversion(off, theta, 5, False, 1e-10, 1)
    assert np.array_equal(o, off) and np.array_equal(t, theta / 5.0)


# ---------------------------------------------------------------------------
# the deck runner's &alter_setup
# ---------------------------------------------------------------------------
def _harm_deck(tmp_path, tail, root):
    import test_deck as T
    deck = T.HARM_IN.replace("&track\nzstop=5\n&end", tail)
    return T._write_planar(tmp_path, deck, root)


def _slice(path, k, i=1):
    return g4mi.h5_read(path, f"slice{i:06d}/{k}")


def test_alter_setup_harmonic_deck_check(tmp_path):
    """&alter_setup harmonic = 3 in a steady-state deck: the 3rd-harmonic
    field becomes the fundamental (same values, wavelength lambda0 / 3, no
    other field), theta is 3 theta, and the next &track writes
    <root>.Run2.out.h5 (Setup::getRootName after incrementCount)."""
    root = str(tmp_path / "c")
    tail = ("&write\nfield = @.pre\nbeam = @.pre\n&end\n\n&alter_setup\nharmonic = 3\n&end\n\n"
            "&write\nfield = @.post\nbeam = @.post\n&end\n\n&track\nzstop=2.5\n&end")
    s = g4mi.check_deck(_harm_deck(tmp_path, tail, root))
    pre3, post = root + ".pre.h3.fld.h5", root + ".post.fld.h5"
    assert not os.path.exists(root + ".post.h3.fld.h5")
    assert g4mi.h5_read(post, "gridpoints")[0] == 41
    assert g4mi.h5_read(post, "wavelength")[0] == pytest.approx(1e-10 / 3, rel=1e-15)
    for k in ("field-real", "field-imag"):
        assert np.array_equal(_slice(post, k), _slice(pre3, k))
    assert np.array_equal(_slice(root + ".post.par.h5", "theta"), _slice(root + ".pre.par.h5", "theta") * 3.0)
    assert np.array_equal(_slice(root + ".post.par.h5", "gamma"), _slice(root + ".pre.par.h5", "gamma"))
    assert os.path.exists(root + ".Run2.out.h5") and not os.path.exists(root + ".out.h5")
    assert s["ngrid"] == 41


def test_alter_setup_subharmonic_deck_check(tmp_path):
    """&alter_setup subharmonic = 3: harmonic fields are deleted and the
    fundamental becomes harmonic 3 of the new reference (its dump
    <name>.h3.fld.h5, wavelength 3 lambda0 / 3); a new &field defines the
    fundamental at 3 lambda0; theta is theta / 3; rootname resets the run
    count."""
    root = str(tmp_path / "c")
    tail = ("&write\nfield = @.pre\nbeam = @.pre\n&end\n\n&alter_setup\nsubharmonic = 3\nrootname = " + root +
            "_sub\n&end\n\n&field\npower=1e3\ndgrid=2e-4\nngrid=35\nwaist_size=40e-6\n&end\n\n"
            "&write\nfield = @.post\nbeam = @.post\n&end\n\n&track\nzstop=2.5\n&end")
    g4mi.check_deck(_harm_deck(tmp_path, tail, root))
    sub = root + "_sub"
    h3 = sub + ".post.h3.fld.h5"
    assert g4mi.h5_read(h3, "wavelength")[0] == pytest.approx(1e-10, rel=1e-15)
    assert g4mi.h5_read(h3, "gridpoints")[0] == 63
    for k in ("field-real", "field-imag"):
        assert np.array_equal(_slice(h3, k), _slice(root + ".pre.fld.h5", k))
    assert g4mi.h5_read(sub + ".post.fld.h5", "gridpoints")[0] == 35
    assert g4mi.h5_read(sub + ".post.fld.h5", "wavelength")[0] == pytest.approx(3e-10, rel=1e-15)
    assert np.array_equal(_slice(sub + ".post.par.h5", "theta"), _slice(root + ".pre.par.h5", "theta") / 3.0)
    assert os.path.exists(sub + ".out.h5")


def test_alter_setup_rejects(tmp_path):
    root = str(tmp_path / "c")
    for body, msg in (("harmonic = 3\ndisable = true", "disable"), ("beamline = NOPE", "lattice"),
                      ("harmonic = 0", "at least 1")):
        tail = f"&alter_setup\n{body}\n&end\n\n&track\nzstop=2.5\n&end"
        with pytest.raises(RuntimeError, match=msg):
            g4mi.check_deck(_harm_deck(tmp_path, tail, root))


def _track_records(sim, st, n, gamma0):
    p = [sim.diag()["power"][0]]
    for i in range(n):
        sim.step(st[i], gamma0)
        p.append(sim.diag()["power"][0])
    return np.array(p)


@pytest.mark.gpu
def test_alter_setup_harmonic_deck_run(tmp_path):
    """Example1 (planar, 63^2) with a seeded 3rd harmonic (41^2), tracked to
    z = 2.5 m; &alter_setup harmonic = 3; tracked again.  The second run's
    /Field/power follows the oracle's run from the oracle's converted state
    (theta 3 theta, the 3rd harmonic as the fundamental at lambda0 / 3), every
    record within 1e-10."""
    print("deck runner: &alter_setup harmonic = 3 between two tracks", flush=True)
    root = str(tmp_path / "c")
    tail = "&track\nzstop=2.5\n&end\n\n&alter_setup\nharmonic = 3\n&end\n\n&track\nzstop=2.5\n&end"
    path = _harm_deck(tmp_path, tail, root)
    s = g4mi.check_deck(path)
    g4mi.run_deck(path, device=0, quiet=True)
    p1 = g4mi.h5_read(root + ".out.h5", "Field/power")[:, 0]
    p2 = g4mi.h5_read(root + ".Run2.out.h5", "Field/power")[:, 0]
    els = []
    for k1 in (2.0, -2.0) * 6:
        els += [("undulator", dict(aw=0.84853, lambdau=0.015, nwig=266, helical=False)),
                ("drift", dict(l=0.44)), ("quadrupole", dict(l=0.08, k1=k1)), ("drift", dict(l=0.24))]
    g0 = 11357.82
    sim = O.Sim([8192], 63, 2e-4, 1e-10, timerun=False, threads=16)
    O.load_beam(sim, dict(gamma=g0, delgam=1.0, current=3000.0, ex=4e-7, ey=4e-7, betax=s["betax"],
                          betay=s["betay"], alphax=s["alphax"], alphay=s["alphay"]), None, 8, None, lambda0=1e-10)
    O.load_gauss_all(sim, 5e3, 30e-6, 1e-10)
    sim.add_harmonics([(3, 41, 1.5e-4)])
    sim.hfields[0][0] = g4mi.load_gauss(41, 1.5e-4, 1e-10 / 3, 1e2, 20e-6)
    r1 = _track_records(sim, O.unroll(els, 0.045, 1e-10, g0), len(p1) - 1, g0)
    # the oracle's conversion: a new state at lambda0 / 3 on the harmonic's grid
    sim2 = O.Sim([8192], 41, 1.5e-4, 1e-10 / 3, timerun=False, threads=16)
    off, theta = O.beam_harmonic_conversion(sim.off, sim.theta, 3, False)
    sim2.theta[:] = theta
    for k in ("gamma", "x", "y", "px", "py", "current"):
        getattr(sim2, k)[:] = getattr(sim, k)
    f, first = O.field_harmonic_conversion(sim.hfields[0], 0, 3, False)
    sim2.field[:] = f
    r2 = _track_records(sim2, O.unroll(els, 0.045, 1e-10 / 3, g0), len(p2) - 1, g0)
    e1 = float(np.max(np.abs(p1 / r1 - 1)))
    e2 = float(np.max(np.abs(p2 / r2 - 1)))
    print(f"  {len(p1)} + {len(p2)} records: P(z) rel err {e1:.3e} / {e2:.3e}, P end {p2[-1]:.4e} W", flush=True)
    assert e1 < 1e-10 and e2 < 1e-10


@pytest.mark.gpu
def test_alter_setup_one4one_resample_deck(tmp_path):
    """A one4one beam (3 slices, sample 2) converted with harmonic = 2 and
    resample = true: the device sort leaves each of the 6 new slices with
    the particles (and currents) of the oracle's conversion and sort; a
    new &field and a &track then run on the 6 slices."""
    import test_one4one as T1
    print("deck runner: one4one harmonic = 2 with resampling", flush=True)
    root = str(tmp_path / "o")
    conv = ("&alter_setup\nharmonic = 2\nresample = true\n&end\n\n&write\nbeam = @.up\n&end\n\n"
            "&field\npower=5e3\ndgrid=2e-4\nngrid=31\nwaist_size=30e-6\n&end\n\n&track")
    path = T1._o4o_deck(tmp_path, root)
    path.write_text(path.read_text().replace("&track", conv, 1))
    s = g4mi.check_deck(path)
    assert s["nslice"] == 6
    g4mi.run_deck(path, device=0, quiet=True)
    sl = [dict(T1.SLICE, current=30.0, delgam=1.0, betax=8.0, betay=16.0, alphax=0.0, alphay=0.0)] * 3
    cnt = O.one4one_counts([30.0] * 3, 1e-10, 2)
    sim = O.Sim(cnt, 31, 2e-4, 1e-10, sample=2)
    O.load_beam_one4one(sim, sl, 9, 1e-10, 2)
    init = root + ".init.par.h5"
    for i in range(3):  # the deck's beam is the oracle's
        a, b = int(sim.off[i]), int(sim.off[i + 1])
        assert np.array_equal(_slice(init, "theta", i + 1), sim.theta[a:b])
    off, theta = O.beam_harmonic_conversion(sim.off, sim.theta, 2, True)
    others = [sim.gamma, sim.x, sim.y, sim.px, sim.py]
    o, th, g, cur = _sort(off, theta, others, 0.5e-10, 2)
    up = root + ".up.par.h5"
    for i in range(6):
        a, b = int(o[i]), int(o[i + 1])
        dt, dg = _slice(up, "theta", i + 1), _slice(up, "gamma", i + 1)
        assert len(dt) == b - a, i
        assert np.array_equal(np.sort(dt), np.sort(th[a:b])), i
        assert np.array_equal(np.sort(dg), np.sort(g[a:b])), i
        assert _slice(up, "current", i + 1)[0] == pytest.approx(cur[i], rel=1e-15)
    assert os.path.exists(root + ".Run2.out.h5")
    print(f"  counts {np.diff(o).tolist()}", flush=True)


def test_importfield_harmonic_round_trip(tmp_path):
    """&importfield harmonic = 3 (ImportField.cpp:42-91, readFieldHDF5.cpp:21-140):
    the 3rd-harmonic dump of a deck comes back as the 3rd-harmonic field of
    another (its wavelength lambda0 / 3 checked against lambda0 / harmonic),
    within one rounding of the sqrt(W) scaling; importing it twice, or before
    the fundamental, is an error."""
    import test_deck as T
    src = str(tmp_path / "s")
    T._write_planar(tmp_path, T.HARM_IN.replace("&track\nzstop=5\n&end", "&write\nfield = @.init\n&end"), src)
    g4mi.check_deck(tmp_path / "Example1.in")
    body = T.EX1_IN.split("&field")[0]  # &setup and &lattice
    imp = (body + f"&importfield\nfile = {src}.init.fld.h5\n&end\n\n"
           f"&importfield\nfile = {src}.init.h3.fld.h5\nharmonic = 3\n&end\n\n"
           "&write\nfield = @.back\n&end\n")
    root = str(tmp_path / "i")
    T._write_planar(tmp_path, imp, root)
    g4mi.check_deck(tmp_path / "Example1.in")
    a, b = src + ".init.h3.fld.h5", root + ".back.h3.fld.h5"
    for k in ("gridsize", "wavelength", "gridpoints"):
        assert np.array_equal(g4mi.h5_read(a, k), g4mi.h5_read(b, k)), k
    for k in ("field-real", "field-imag"):
        va, vb = _slice(a, k), _slice(b, k)
        assert np.max(np.abs(va - vb)) <= 4e-16 * np.max(np.abs(va)), k
    twice = imp.replace("&write", f"&importfield\nfile = {src}.init.h3.fld.h5\nharmonic = 3\n&end\n\n&write")
    T._write_planar(tmp_path, twice, root)
    with pytest.raises(RuntimeError, match="already defined"):
        g4mi.check_deck(tmp_path / "Example1.in")
    first = body + f"&importfield\nfile = {src}.init.h3.fld.h5\nharmonic = 3\n&end\n"
    T._write_planar(tmp_path, first, root)
    with pytest.raises(RuntimeError, match="fundamental"):
        g4mi.check_deck(tmp_path / "Example1.in")


@pytest.mark.gpu
def test_restart_with_harmonic_field(tmp_path):
    """Checkpoint and restart with a 3rd-harmonic field: the planar Example1
    (63^2, harmonic 41^2) dumps beam and both fields after one FODO cell and
    runs a second cell; a second deck imports the beam, the fundamental and
    (&importfield harmonic = 3) the harmonic dump and tracks one cell.  Its
    /Field/power and /Field3/power are the uninterrupted run's from the
    dump on, within 1e-10."""
    import test_deck as T
    print("deck runner: restart from dumps with a 3rd-harmonic field", flush=True)
    base = T.HARM_IN.replace("&track\nzstop=5\n&end", "&track\n&end")
    path = T._write_planar(tmp_path, base, str(tmp_path / "full"))
    s = g4mi.check_deck(path)
    k = s["steps"] // 6
    z = g4mi.h5_read(str(tmp_path / "full.out.h5"), "Lattice/z")
    deck_a = base.replace("&track\n&end", f"&track\nzstop = {z[2 * k] - 1e-3}\nfield_dump_step = {k}\n"
                          f"beam_dump_step = {k}\n&end")
    path = T._write_planar(tmp_path, deck_a, str(tmp_path / "full"))
    g4mi.run_deck(path, device=0, quiet=True)
    out_a = str(tmp_path / "full.out.h5")
    pa, p3a = g4mi.h5_read(out_a, "Field/power")[:, 0], g4mi.h5_read(out_a, "Field3/power")[:, 0]
    src = str(tmp_path / f"full.{k}")
    assert os.path.exists(src + ".h3.fld.h5")
    deck_b = (base.replace("&field\npower=5e3\ndgrid=2.000000e-04\nngrid=63\nwaist_size=30e-6\n&end",
                           f"&importfield\nfile = {src}.fld.h5\n&end")
              .replace("&field\nharm=3\npower=1e2\ndgrid=1.5e-4\nngrid=41\nwaist_size=20e-6\n&end",
                       f"&importfield\nfile = {src}.h3.fld.h5\nharmonic = 3\n&end")
              .replace("&beam\ncurrent=3000\ndelgam=1.000000\nex=4.000000e-07\ney=4.000000e-07\n&end",
                       f"&importbeam\nfile = {src}.par.h5\n&end")
              .replace("&track\n&end", f"&track\nzstop = {z[k] - 1e-3}\n&end"))
    assert deck_b.count("&importfield") == 2 and "&importbeam" in deck_b
    path = T._write_planar(tmp_path, deck_b, str(tmp_path / "rest"))
    g4mi.run_deck(path, device=0, quiet=True)
    out_b = str(tmp_path / "rest.out.h5")
    pb, p3b = g4mi.h5_read(out_b, "Field/power")[:, 0], g4mi.h5_read(out_b, "Field3/power")[:, 0]
    assert len(pb) == k + 1
    ep = float(np.max(np.abs(pb / pa[k:2 * k + 1] - 1)))
    e3 = float(np.max(np.abs(p3b / p3a[k:2 * k + 1] - 1)))
    print(f"  restart after {k} steps: P(z) rel err {ep:.3e}, P3(z) {e3:.3e}", flush=True)
    assert ep < 1e-10 and e3 < 1e-10
