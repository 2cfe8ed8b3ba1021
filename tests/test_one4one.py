"""one4one beams: loading from one ran2 stream (QuietLoading with
one4one) and Beam::sort's migration of particles between slices.

CPU: the oracle's loader against the reference's own QuietLoading
(oracle/_ref, bit for bit), the library's host loader against the oracle
(bit for bit), the oracle's sort against a numpy brute force.
GPU: the device sort against the oracle's (the same particles in every
slice, bit for bit; the order within a slice differs), and a SASE run with
sort markers against the oracle, P(z) within 1e-10."""
import ctypes as C
import math
import os
import sys

import numpy as np
import pytest

import cases
import g4mi
import harness as H
import oracle_py as O

LAM = cases.LAMBDA_R
SLICE = dict(gamma=11357.82, delgam=2.0, current=50.0, ex=4e-7, ey=4e-7, betax=10.0, betay=12.0,
             alphax=0.3, alphay=-0.2)


def _slices(ns, cur0=50.0):
    return [dict(SLICE, current=cur0 * (1 + 0.1 * j)) for j in range(ns)]


def _oracle_beam(ns, seed=7, sample=1, ngrid=31, gridmax=0.2e-3, timerun=True):
    sl = _slices(ns)
    cnt = O.one4one_counts([s["current"] for s in sl], LAM, sample)
    sim = O.Sim(cnt, ngrid, gridmax, LAM, sample=sample, timerun=timerun)
    O.load_beam_one4one(sim, sl, seed, LAM, sample)
    return sim, sl, cnt


def test_one4one_counts_and_loader_match_oracle():
    sim, sl, cnt = _oracle_beam(5)
    assert list(cnt) == [round(s["current"] * LAM / O.CE) for s in sl]
    assert list(g4mi.one4one_counts(sl, LAM)) == list(cnt)
    b = g4mi.load_beam_one4one(sim.off, sl, 7)
    for k in ("gamma", "theta", "x", "y", "px", "py"):
        assert np.array_equal(b[k], getattr(sim, k)), k


def test_one4one_loader_matches_reference_build():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    sim, sl, cnt = _oracle_beam(4, seed=31, sample=2)
    sp = np.array([[s.get(k, 0.0) for k in (
        "gamma", "delgam", "current", "ex", "ey", "xcen", "ycen", "pxcen", "pycen",
        "betax", "betay", "alphax", "alphay", "bunch", "bunchphase", "emod", "emodphase")] for s in sl])
    out = {k: np.zeros(len(sim.gamma)) for k in ("gamma", "theta", "x", "y", "px", "py")}
    R.ref_quiet_load_one4one(31, 0, 4, sim.off.ctypes.data_as(C.POINTER(C.c_longlong)),
                             np.ascontiguousarray(sp).ctypes.data_as(C.POINTER(C.c_double)),
                             4 * math.asin(1) * 2, *[out[k].ctypes.data_as(C.POINTER(C.c_double)) for k in
                                                    ("gamma", "theta", "x", "y", "px", "py")])
    for k in out:
        assert np.array_equal(out[k], getattr(sim, k)), k


def _brute_sort(sim, slen):
    """Per slice, the sorted list of particles (theta, gamma, x, y, px, py)
    after migration, computed particle by particle."""
    ns = sim.ns
    dest = [[] for _ in range(ns)]
    for a in range(ns):
        for i in range(int(sim.off[a]), int(sim.off[a + 1])):
            th = sim.theta[i]
            s = slen * a + th
            if s < 0 or s > slen * ns:
                continue
            atar = math.floor(th * (1.0 / slen))
            if a + atar >= ns or a + atar < 0:
                continue
            dest[a + atar].append((th - slen * atar, sim.gamma[i], sim.x[i], sim.y[i], sim.px[i], sim.py[i]))
    return [sorted(d) for d in dest]


def _per_slice(off, b):
    return [sorted(zip(*[b[k][int(off[a]):int(off[a + 1])] for k in ("theta", "gamma", "x", "y", "px", "py")]))
            for a in range(len(off) - 1)]


def _scramble(sim, rng, slen, spread=1.6):
    """theta moved by up to +-spread slices, so particles leave their slice
    in both directions and some leave the window."""
    sim.theta += rng.uniform(-spread, spread, size=len(sim.theta)) * slen


def _plant_lower_edge(sim, slen):
    """Puts the first particle of a slice a > 0 at s = slen a + theta = 0
    exactly, in a slice where floor(theta / slen) rounds to -a - 1, one
    below the first slice (slen is the sort's own, 4 asin(1) slicelength /
    reflength; at sample 9 this happens in slice 5).  Returns the index."""
    assert slen == 4.0 * np.arcsin(1.0) * sim.s.slicelength / sim.s.reflength
    for a in range(1, sim.ns):
        th = -(slen * a)
        if slen * a + th == 0.0 and math.floor(th * (1.0 / slen)) == -a - 1 and sim.off[a + 1] > sim.off[a]:
            i = int(sim.off[a])
            sim.theta[i] = th
            return i
    raise AssertionError("no slice rounds below the first one at this slice length")


def test_oracle_sort_drops_lower_edge_particle():
    sim, sl, cnt = _oracle_beam(8, sample=9, ngrid=15)
    slen = 4.0 * np.arcsin(1.0) * sim.s.slicelength / sim.s.reflength
    _plant_lower_edge(sim, slen)
    ref = _brute_sort(sim, slen)
    n0 = int(sim.off[-1])
    sim.sort()
    got = _per_slice(sim.off, {k: getattr(sim, k) for k in ("gamma", "theta", "x", "y", "px", "py")})
    assert got == ref
    assert int(sim.off[-1]) == n0 - 1


def test_oracle_sort_matches_brute_force():
    sim, sl, cnt = _oracle_beam(6, sample=2)
    slen = 2 * math.pi * 2
    _scramble(sim, np.random.default_rng(3), slen)
    ref = _brute_sort(sim, slen)
    sim.sort()
    got = _per_slice(sim.off, {k: getattr(sim, k) for k in ("gamma", "theta", "x", "y", "px", "py")})
    assert got == ref
    assert np.array_equal(sim.current, np.diff(sim.off) * O.CE / (2 * LAM))
    assert all(0 <= p[0] < slen for sl_ in got for p in sl_)
    assert sum(len(g) for g in got) < int(sum(cnt))  # some left the window


@pytest.mark.gpu
def test_device_sort_matches_oracle():
    print("one4one sort: 8 slices, scrambled phases, GPU vs oracle", flush=True)
    sim, sl, cnt = _oracle_beam(8, sample=9, ngrid=15)
    slen = 4.0 * np.arcsin(1.0) * sim.s.slicelength / sim.s.reflength
    _scramble(sim, np.random.default_rng(11), slen)
    _plant_lower_edge(sim, slen)
    ctx = H.to_gpu(sim)
    ctx.set_sorting(True)
    n = ctx.sort()
    sim.sort()
    off, cur = ctx.slices()
    assert n == int(sim.off[-1]) and np.array_equal(off, sim.off)
    assert np.array_equal(cur, sim.current)
    b = ctx.download_beam()
    assert _per_slice(off, b) == _per_slice(sim.off, {k: getattr(sim, k) for k in b})
    # a second sort moves nothing
    assert ctx.sort() == n and np.array_equal(ctx.slices()[0], off)
    ctx.close()


def _stable_order(sim, slen):
    """The device sort's documented result: particles keyed by destination
    slice (dropped ones last), each slice in source order, theta shifted by
    slen atar; as arrays and offsets."""
    ns = sim.ns
    a = np.repeat(np.arange(ns), np.diff(sim.off))
    th = sim.theta
    s = slen * a + th
    atar = np.floor(th * (1.0 / slen))
    dest = a + atar
    ok = ~((s < 0) | (s > slen * ns)) & (dest >= 0) & (dest < ns)
    key = np.where(ok, dest, ns).astype(np.int64)
    th_new = np.where(ok, th - slen * atar, th)
    order = np.argsort(key, kind="stable")
    keep = order[key[order] < ns]
    out = {k: getattr(sim, k)[keep] for k in ("gamma", "x", "y", "px", "py")}
    out["theta"] = th_new[keep]
    off = np.concatenate([[0], np.cumsum(np.bincount(key[keep], minlength=ns))]).astype(np.int64)
    return out, off


@pytest.mark.gpu
@pytest.mark.parametrize("ns,sample", [(40, 9), (300, 1)])
def test_device_sort_is_stable(ns, sample):
    """The hand-written radix sort over many tiles of 2048 keys and two
    (ns = 40) or three (ns = 300) 4-bit digit passes: every slice holds its
    particles in source order, bit for bit as a stable sort by destination."""
    print(f"one4one sort order: {ns} slices, scrambled phases", flush=True)
    sim, sl, cnt = _oracle_beam(ns, sample=sample, ngrid=15)
    slen = 4.0 * np.arcsin(1.0) * sim.s.slicelength / sim.s.reflength
    _scramble(sim, np.random.default_rng(5), slen, spread=3.5)
    ref, roff = _stable_order(sim, slen)
    ctx = H.to_gpu(sim)
    ctx.set_sorting(True)
    n = ctx.sort()
    off, _ = ctx.slices()
    assert n == int(roff[-1]) and np.array_equal(off, roff)
    b = ctx.download_beam()
    for k in ref:
        assert np.array_equal(b[k], ref[k]), k
    ctx.close()


@pytest.mark.gpu
def test_one4one_sase_with_sort_markers():
    """A one4one SASE run (8 slices, 1.3-1.6 k electrons each, 31^2) with a
    sort every 10 steps over 150 steps: P(z) and |b(z)| against the oracle
    within 1e-10; the particle counts per slice after the run are equal."""
    print("one4one SASE with sorting, 150 steps, GPU vs oracle", flush=True)
    st = H.steps()[:150]
    for i, s in enumerate(st):
        if i % 10 == 0:
            s.marker |= 4
    sim, sl, cnt = _oracle_beam(8, seed=3, ngrid=31, gridmax=0.2e-3)
    sim.s.ndiag_threads = 16
    ctx = H.to_gpu(sim)
    ctx.set_sorting(True)
    g = ctx.run(H.gsteps(st), cases.GAMMA0, record=("power", "bunching"))
    r = sim.run(st, cases.GAMMA0, record=("power", "bunching"))
    ep = H.rel_err(g["power"][1:], r["power"][1:])
    eb = H.rel_err(g["bunching"], r["bunching"])
    off, _ = ctx.slices()
    print(f"  P(z) rel err {ep:.3e}  |b(z)| {eb:.3e}  particles {int(sum(cnt))} -> {int(off[-1])}", flush=True)
    assert np.array_equal(off, sim.off)
    assert ep < 1e-10 and eb < 1e-10
    ctx.close()


O4O_IN = """&setup
rootname={root}
lattice={lat}
beamline=FEL
lambda0=1e-10
gamma0=11357.82
delz=0.045000
one4one=true
seed=9
&end

&time
slen=6e-10
sample=2
&end

&field
power=5e3
dgrid=2.000000e-04
ngrid=31
waist_size=30e-6
&end

&beam
current=30
delgam=1.000000
ex=4.000000e-07
ey=4.000000e-07
betax=8
betay=16
&end

&write
beam = @.init
&end

&track
zstop=3
sort_step=10
beam_dump_step=40
{extra}
&end
"""


def _o4o_deck(tmp_path, root, extra=""):
    import test_deck as T
    (tmp_path / "Example1.lat").write_text(T.EX1_LAT.format(extra=""))
    path = tmp_path / "o4o.in"
    path.write_text(O4O_IN.format(root=root, lat=str(tmp_path / "Example1.lat"), extra=extra))
    return path


def test_one4one_deck_setup(tmp_path):
    """&setup one4one = true: electrons per slice round(I lambda0 sample /
    ce), loaded as the oracle's one4one loader; the .par.h5 dump says
    one4one = 1 with beamlets of 1."""
    root = str(tmp_path / "o")
    s = g4mi.check_deck(_o4o_deck(tmp_path, root))
    ne = round(30 * 1e-10 * 2 / O.CE)
    assert s["nslice"] == 3 and s["npart"] == 3 * ne
    par = root + ".init.par.h5"
    assert g4mi.h5_read(par, "one4one")[0] == 1 and g4mi.h5_read(par, "beamletsize")[0] == 1
    sl = [dict(SLICE, current=30.0, delgam=1.0, betax=8.0, betay=16.0, alphax=0.0, alphay=0.0)] * 3
    cnt = O.one4one_counts([30.0] * 3, 1e-10, 2)
    sim = O.Sim(cnt, 31, 2e-4, 1e-10, sample=2)
    O.load_beam_one4one(sim, sl, 9, 1e-10, 2)
    for i in range(3):
        a, b = int(sim.off[i]), int(sim.off[i + 1])
        for k in ("gamma", "theta", "x", "y", "px", "py"):
            assert np.array_equal(g4mi.h5_read(par, f"slice{i + 1:06d}/{k}"), getattr(sim, k)[a:b]), (i, k)


@pytest.mark.gpu
def test_one4one_deck_run(tmp_path):
    """The one4one deck on the device: sorts every 10 steps, beam dumps hold
    the sorted slices (their counts add up to the surviving particles) and
    the slice currents follow np ce / slicelength."""
    print("deck runner: one4one run with sorting and beam dumps", flush=True)
    root = str(tmp_path / "o")
    s = g4mi.check_deck(_o4o_deck(tmp_path, root))
    g4mi.run_deck(_o4o_deck(tmp_path, root), device=0, quiet=True)
    steps = list(g4mi.h5_read(root + ".out.h5", "Meta/Beamdumps/intstep"))
    assert steps[0] == 0 and len(steps) >= 2
    last = f"{root}.{int(steps[-1])}.par.h5"
    n = [len(g4mi.h5_read(last, f"slice{i + 1:06d}/gamma")) for i in range(3)]
    cur = [g4mi.h5_read(last, f"slice{i + 1:06d}/current")[0] for i in range(3)]
    assert np.allclose(cur, np.array(n) * O.CE / (2 * 1e-10), rtol=1e-15)
    assert sum(n) <= s["npart"]
