"""CPU checks of libg4mi's host side: exported C-ABI, input generators
(beam, field, lattice) bit-exact against the oracle, and error handling.
No GPU is used here."""
import ctypes as C
import re

import numpy as np
import pytest

import cases
import g4mi
import oracle_py as O


def test_library_exports_every_declared_symbol():
    text = open(g4mi.HEADER).read()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(g4mi_\w+)\s*\(", text, re.M))
    assert len(declared) >= 20
    L = g4mi.lib()
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing


def _slices(ns, currents=None):
    cur = currents if currents is not None else [cases.CURRENT] * ns
    return [cases.slice_beam(current=c) for c in cur]


@pytest.mark.parametrize("counts,shot", [([1024] * 6, True), ([256, 4096, 512, 1024], True),
                                         ([2048], False)])
def test_beam_loader_bit_exact(counts, shot):
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    ns = len(counts)
    cur = [cases.CURRENT * (1 + 0.1 * i) for i in range(ns)]
    got = g4mi.load_beam(off, _slices(ns, cur), 4, shot, 7, cases.LAMBDA_R, 1.0, True)
    sim = O.Sim(counts, 11, cases.GRIDMAX, cases.LAMBDA_R)
    O.load_beam(sim, _slices(ns, cur), None, 4, 7 if shot else None, lambda0=cases.LAMBDA_R)
    for k in ("gamma", "theta", "x", "y", "px", "py"):
        assert np.array_equal(got[k], getattr(sim, k)), k


def test_beam_loader_rejects_bad_nbins():
    off = np.array([0, 1022], np.int64)
    with pytest.raises(g4mi.G4miError, match="NBINS"):
        g4mi.load_beam(off, _slices(1), 4, False, 1, cases.LAMBDA_R)


def test_gauss_field_bit_exact():
    f = g4mi.load_gauss(cases.NGRID, cases.GRIDMAX, cases.LAMBDA_R, 5e3, 30e-6, 0.7, 0.3)
    ref = np.zeros(cases.NGRID ** 2 * 2)
    O.lib().orc_load_gauss(ref.ctypes.data_as(C.POINTER(C.c_double)), cases.NGRID,
                           cases.GRIDMAX, cases.LAMBDA_R, 5e3, 30e-6, 0.7, 0.3, 0.0, 0.0)
    assert np.array_equal(f.reshape(-1), ref)


def _assert_steps_equal(a, b):
    assert len(a) == len(b)
    for i, (s, t) in enumerate(zip(a, b)):
        for k in g4mi.STEP_FIELDS + ("helical", "marker"):
            assert getattr(s, k) == getattr(t, k), (i, k, getattr(s, k), getattr(t, k))


def test_lattice_parse_matches_oracle_unroll():
    got = g4mi.parse_lattice(cases.lattice_text(), "FEL", cases.DELZ, cases.LAMBDA_R,
                             cases.GAMMA0)
    ref = O.unroll(cases.fodo_elements(), cases.DELZ, cases.LAMBDA_R, cases.GAMMA0)
    _assert_steps_equal(got, ref)
    # 4 x 125 undulator steps + 3 x (drift, quad, drift)
    assert len(got) == 509
    # autophasing slippage after each drift section (Lattice.cpp:171-177)
    slips = np.array([s.slip for s in got])
    assert np.all(np.abs(slips[:125] - 1) < 1e-9)
    drift_end = 125 + 3 - 1
    assert slips[drift_end] == np.floor(0.5 / 2 / cases.GAMMA0 ** 2 / cases.LAMBDA_R) + 1


def test_lattice_example1_helical_and_markers():
    text = """# Example1 beamline with a marker
D1: DRIFT = { l = 0.44};
D2: DRIFT = { l = 0.24};
QF: QUADRUPOLE = { l = 0.080000, k1= 2.000000 };
QD: QUADRUPOLE = { l = 0.080000, k1= -2.000000 };
UND: UNDULATOR = { lambdau=0.015000, nwig=266, aw=0.84853, helical= True};
M: MARKER = {dumpfield=1, stop=0};
FODO: LINE={UND,D1,QF,D2,M,UND,D1,QD,D2};
FEL: LINE={6*FODO};
"""
    got = g4mi.parse_lattice(text, "fel", 0.045, 1e-10, 11357.82)
    und = ("undulator", dict(aw=0.84853, lambdau=0.015, nwig=266, helical=True))
    fodo = [und, ("drift", dict(l=0.44)), ("quadrupole", dict(l=0.08, k1=2.0)),
            ("drift", dict(l=0.24)), ("marker", dict(action=1)), und, ("drift", dict(l=0.44)),
            ("quadrupole", dict(l=0.08, k1=-2.0)), ("drift", dict(l=0.24))]
    ref = O.unroll(fodo * 6, 0.045, 1e-10, 11357.82)
    _assert_steps_equal(got, ref)
    assert sum(1 for s in got if s.marker & 1) == 6


def test_lattice_errors_are_reported():
    with pytest.raises(g4mi.G4miError, match="beamline"):
        g4mi.parse_lattice(cases.lattice_text(), "NOPE", 0.015, 1e-10, 1e4)
    with pytest.raises(g4mi.G4miError, match="undefined"):
        g4mi.parse_lattice("A: LINE={X};", "A", 0.015, 1e-10, 1e4)


def test_create_rejects_bad_arguments():
    h = C.c_void_p()
    off = np.array([0, 10], np.int64)
    rc = g4mi.lib().g4mi_create(C.byref(h), 0, 1, off.ctypes.data_as(C.POINTER(C.c_longlong)),
                                2, 1e-3, 1e-10, 1.0, 1, 0, 0)
    assert rc == -1 and b"invalid" in g4mi.lib().g4mi_last_error()
    rc = g4mi.lib().g4mi_create(C.byref(h), 0, 1, off.ctypes.data_as(C.POINTER(C.c_longlong)),
                                11, 1e-3, 1e-10, 1.0, 1, 0, 7)
    assert rc == -1


def test_create_without_gpu_fails_loudly():
    h = C.c_void_p()
    off = np.array([0, 8], np.int64)
    rc = g4mi.lib().g4mi_create(C.byref(h), 0, 1, off.ctypes.data_as(C.POINTER(C.c_longlong)),
                                11, 1e-3, 1e-10, 1.0, 1, 0, 0)
    if rc == 0:   # a GPU is present: the context works and is released
        g4mi.lib().g4mi_destroy(h)
    else:
        assert rc == -3 and b"device" in g4mi.lib().g4mi_last_error()
