"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; it is the checker, never the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libg4oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libg4ref.so")

# Physical constants of the reference (src/Main/GenMain.cpp:63-68).
VACIMP = 376.73
EEV = 510998.95069
CE = 4.8032045e-11

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_longlong)


class Step(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "aw", "ax", "ay", "ku", "kx", "ky", "gradx", "grady", "qf", "qx", "qy", "cx", "cy",
        "chic_angle", "chic_lb", "chic_ld", "chic_lt", "slip", "phaseshift", "z", "dz")] + [
        ("helical", C.c_int), ("marker", C.c_int)]


class HField(C.Structure):
    _fields_ = [("harm", C.c_int), ("ngrid", C.c_int), ("gridmax", C.c_double), ("dgrid", C.c_double),
                ("xks", C.c_double), ("field", C.POINTER(C.c_double))]


class State(C.Structure):
    _fields_ = [
        ("nslice", C.c_int), ("off", _lp),
        ("gamma", _dp), ("theta", _dp), ("x", _dp), ("y", _dp), ("px", _dp), ("py", _dp),
        ("current", _dp), ("longESC", _dp), ("eloss", _dp),
        ("ngrid", C.c_int), ("gridmax", C.c_double), ("dgrid", C.c_double), ("xks", C.c_double),
        ("xlambda", C.c_double), ("slicelength", C.c_double), ("field", _dp),
        ("first", C.c_int), ("accuslip", C.c_double),
        ("sc_nz", C.c_int), ("sc_nphi", C.c_int), ("sc_ngrid", C.c_int), ("longrange", C.c_int),
        ("sc_rmax", C.c_double), ("sc_ks", C.c_double), ("sc_efield", _dp), ("sc_ez", _dp),
        ("reflength", C.c_double), ("sample", C.c_double), ("timerun", C.c_int),
        ("periodic", C.c_int), ("ndiag_threads", C.c_int),
        ("fft_solver", C.c_int), ("fft_filter", C.c_int),
        ("fft_xcut", C.c_double), ("fft_ycut", C.c_double), ("fft_sig", C.c_double),
        ("incoh", C.c_void_p), ("nbins", C.c_int), ("wake", _dp),
        ("nh", C.c_int), ("hf", C.POINTER(HField))]


class SliceParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "gamma", "delgam", "current", "ex", "ey", "xcen", "ycen", "pxcen", "pycen",
        "betax", "betay", "alphax", "alphay", "bunch", "bunchphase", "emod", "emodphase")]


DIAG_FIELDS = ("power", "bunching", "bunchingphase", "energy", "energyspread", "xposition",
               "yposition", "xsize", "ysize", "f_xposition", "f_yposition", "f_xsize",
               "f_ysize", "nf_intensity", "ff_intensity", "global_energy")


class DiagOut(C.Structure):
    _fields_ = [(n, _dp) for n in DIAG_FIELDS]


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run make -C oracle")
        L = C.CDLL(LIB_PATH)
        L.orc_ran2.argtypes = [C.c_uint, C.c_int, _dp]
        L.orc_hammersley.argtypes = [C.c_int, C.c_uint, C.c_int, _dp]
        L.orc_inverfc.argtypes = [C.c_double]
        L.orc_inverfc.restype = C.c_double
        L.orc_besselj.argtypes = [C.c_int, C.c_double]
        L.orc_besselj.restype = C.c_double
        L.orc_fc.argtypes = [C.POINTER(Step), C.c_int]
        L.orc_fc.restype = C.c_double
        L.orc_quiet_load.argtypes = [_dp] * 6 + [C.POINTER(SliceParams), C.c_int, C.c_int,
                                                  C.c_double, C.c_int]
        L.orc_shotnoise_new.argtypes = [C.c_int, C.c_int]
        L.orc_shotnoise_new.restype = C.c_void_p
        L.orc_shotnoise_free.argtypes = [C.c_void_p]
        L.orc_quiet_load_one4one.argtypes = [C.c_int, C.c_int, C.c_int, _lp, C.POINTER(SliceParams),
                                             C.c_double] + [_dp] * 6
        L.orc_sort_one4one.argtypes = [C.c_int, _lp, C.c_longlong, C.c_double, C.c_double] + [_dp] * 7
        L.orc_sort_one4one.restype = C.c_longlong
        L.orc_diag_hpower.argtypes = [C.POINTER(State), C.c_int, _dp]
        L.orc_incoherent_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_incoherent_new.restype = C.c_void_p
        L.orc_incoherent_free.argtypes = [C.c_void_p]
        L.orc_incoherent_apply.argtypes = [C.POINTER(State), C.POINTER(Step), C.c_double, C.c_double]
        L.orc_shotnoise.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_double]
        L.orc_load_gauss.argtypes = [_dp, C.c_int] + [C.c_double] * 8
        L.orc_unroll.argtypes = [C.c_int, _ip, _dp, _dp, _dp, C.c_double, C.c_double,
                                 C.c_double, C.POINTER(Step), C.c_int]
        L.orc_unroll.restype = C.c_int
        L.orc_track_transverse.argtypes = [C.POINTER(State), C.POINTER(Step), C.c_double,
                                           C.c_double, C.c_int]
        L.orc_lsc_long.argtypes = [C.POINTER(State), C.c_double, C.c_double]
        L.orc_advance.argtypes = [C.POINTER(State), C.POINTER(Step), C.c_double, C.c_double]
        L.orc_adi_factors.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                      _dp, _dp]
        L.orc_deposit_slice.argtypes = [C.POINTER(State), C.POINTER(Step), C.c_int,
                                        C.c_double, _dp]
        L.orc_adi_slice.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp]
        L.orc_field_advance.argtypes = [C.POINTER(State), C.POINTER(Step), C.c_double]
        L.orc_slippage.argtypes = [C.POINTER(State), C.c_double]
        L.orc_beam_track.argtypes = [C.POINTER(State), C.POINTER(Step), C.c_double, C.c_double]
        L.orc_step_all.argtypes = [C.POINTER(State), C.POINTER(Step), C.c_double]
        L.orc_diag.argtypes = [C.POINTER(State), C.POINTER(DiagOut)]
        L.orc_sort_index.argtypes = [_dp, C.c_int, C.c_double, _ip]
        L.orc_sort_index.restype = C.c_int
        _lib = L
    return _lib


def ref():
    """The reference's own leaf loader chain (oracle/_ref), or None if absent."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        R = C.CDLL(REF_PATH)
        R.ref_ran2.argtypes = [C.c_uint, C.c_int, _dp]
        R.ref_hammersley.argtypes = [C.c_int, C.c_uint, C.c_int, _dp]
        R.ref_inverfc.argtypes = [C.c_double]
        R.ref_inverfc.restype = C.c_double
        R.ref_besselj.argtypes = [C.c_int, C.c_double]
        R.ref_besselj.restype = C.c_double
        R.ref_quiet_load.argtypes = [_dp] * 7 + [C.c_int, C.c_int, C.c_double, C.c_int]
        R.ref_quiet_load_one4one.argtypes = [C.c_int, C.c_int, C.c_int, _lp, _dp, C.c_double] + [_dp] * 6
        R.ref_shotnoise_new.argtypes = [C.c_int, C.c_int]
        R.ref_shotnoise_new.restype = C.c_void_p
        R.ref_shotnoise_free.argtypes = [C.c_void_p]
        R.ref_shotnoise.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_double]
        R.ref_load_gauss.argtypes = [_dp, C.c_int] + [C.c_double] * 8
        _ref = R
    return _ref


# ---------------------------------------------------------------- lattice
ELEM = {"undulator": 0, "drift": 1, "quadrupole": 2, "corrector": 3, "chicane": 4,
        "marker": 5, "phaseshifter": 6}


def layout(elements):
    """Lay out a beamline of (kind, params) back to back as the lattice
    parser does (z = previous z + previous length).  params for undulator:
    dict(aw, lambdau, nwig, kx?, ky?, helical?); drift: l; quadrupole: l, k1.
    Returns arrays (type, z, l, par[8]) for orc_unroll."""
    types, zs, ls, pars = [], [], [], []
    z = 0.0
    for kind, p in elements:
        par = [0.0] * 8
        if kind == "undulator":
            helical = bool(p.get("helical", False))
            kx = p.get("kx", 0.5 if helical else 0.0)
            ky = p.get("ky", 0.5 if helical else 1.0)
            length = p["lambdau"] * p["nwig"]
            par = [p["aw"], p["lambdau"], p["nwig"], kx, ky, p.get("ax", 0.0),
                   p.get("ay", 0.0), 1.0 if helical else 0.0]
        elif kind == "quadrupole":
            length = p["l"]
            par[0:3] = [p["k1"], p.get("dx", 0.0), p.get("dy", 0.0)]
        elif kind == "drift":
            length = p["l"]
        elif kind == "corrector":
            length = p.get("l", 0.0)
            par[0:2] = [p.get("cx", 0.0), p.get("cy", 0.0)]
        elif kind == "chicane":
            length = p["l"]
            par[0:3] = [p["delay"], p["lb"], p["ld"]]
        elif kind == "marker":
            length = 0.0
            par[0] = p.get("action", 0)
        elif kind == "phaseshifter":
            length = p.get("l", 0.0)
            par[0] = p.get("phi", 0.0)
        else:
            raise ValueError(kind)
        types.append(ELEM[kind])
        zs.append(z)
        ls.append(length)
        pars.extend(par)
        z = z + length + 0.0
    return (np.array(types, np.int32), np.array(zs), np.array(ls), np.array(pars))


def unroll(elements, delz, lambda0, gamma0, max_steps=100000):
    t, z, l, par = layout(elements)
    out = (Step * max_steps)()
    n = lib().orc_unroll(len(t), t.ctypes.data_as(_ip), _ptr(z), _ptr(l), _ptr(par), delz,
                         lambda0, gamma0, out, max_steps)
    if n < 0:
        raise RuntimeError("lattice too long")
    return (Step * n)(*out[:n])


def fodo_sase(nund=4, nwig=125, lambdau=0.015, aw=1.7, drift=0.5, lquad=0.08, k1=2.0,
              helical=False):
    """Four planar undulator modules with FODO drift sections between them:
    UND, D(drift-lquad)/2, Q, D(drift-lquad)/2 alternating focusing sign.
    The last module is followed by nothing."""
    els = []
    half = 0.5 * (drift - lquad)
    for i in range(nund):
        els.append(("undulator", dict(aw=aw, lambdau=lambdau, nwig=nwig, helical=helical)))
        if i < nund - 1:
            els.append(("drift", dict(l=half)))
            els.append(("quadrupole", dict(l=lquad, k1=k1 if i % 2 == 0 else -k1)))
            els.append(("drift", dict(l=half)))
    return els


# ---------------------------------------------------------------- state
class Sim:
    """Owns numpy arrays for an oracle run and a matching State struct."""

    def __init__(self, counts, ngrid, gridmax, lambda0, sample=1, timerun=True,
                 periodic=False, sc=(0, 0, 100, 0.0), longrange=False, threads=1):
        counts = np.asarray(counts, dtype=np.int64)
        self.ns = len(counts)
        self.off = np.zeros(self.ns + 1, np.int64)
        self.off[1:] = np.cumsum(counts)
        n = int(self.off[-1])
        self.gamma, self.theta, self.x, self.y, self.px, self.py = (np.zeros(n) for _ in range(6))
        self.current = np.zeros(self.ns)
        self.longESC = np.zeros(self.ns)
        self.eloss = np.zeros(self.ns)
        self.field = np.zeros((self.ns, ngrid * ngrid, 2))
        self.sc_efield = np.zeros(self.ns)
        self.sc_ez = np.zeros(max(1, int(counts.max()) if len(counts) else 1))
        s = State()
        s.nslice = self.ns
        s.off = self.off.ctypes.data_as(_lp)
        for k in ("gamma", "theta", "x", "y", "px", "py", "current", "longESC", "eloss",
                  "field", "sc_efield", "sc_ez"):
            setattr(s, k, _ptr(getattr(self, k)))
        s.ngrid = ngrid
        s.gridmax = gridmax
        s.dgrid = 2 * gridmax / float(ngrid - 1)
        s.xlambda = lambda0 / 1.0
        s.xks = 4.0 * np.arcsin(1.0) / s.xlambda
        s.slicelength = sample * lambda0
        s.first = 0
        s.accuslip = 0.0
        s.sc_nz, s.sc_nphi, s.sc_ngrid, s.sc_rmax = sc[0], sc[1], sc[2], sc[3]
        s.longrange = 1 if longrange else 0
        s.sc_ks = 4.0 * np.arcsin(1.0) / lambda0
        s.reflength = lambda0
        s.sample = float(sample)
        s.timerun = 1 if timerun else 0
        s.periodic = 1 if periodic else 0
        s.ndiag_threads = threads
        self.s = s

    def diag(self):
        out = {k: np.zeros(self.ns) for k in DIAG_FIELDS}
        d = DiagOut(*[_ptr(out[k]) for k in DIAG_FIELDS])
        lib().orc_diag(C.byref(self.s), C.byref(d))
        out["global_energy"] = out["global_energy"][0]
        return out

    one4one = False  # Beam::sort at sort markers (one4one beams)

    def add_harmonics(self, specs):
        """Fields at harmonics besides the fundamental: specs = list of
        (harm, ngrid, gridmax); arrays in self.hfields, zero-filled."""
        self.hfields = [np.zeros((self.ns, ng * ng, 2)) for (_, ng, _) in specs]
        arr = (HField * len(specs))()
        for k, (h, ng, gm) in enumerate(specs):
            arr[k] = HField(h, ng, gm, 2 * gm / float(ng - 1), h * 4.0 * np.arcsin(1.0) / self.s.xlambda,
                            _ptr(self.hfields[k]))
        self._harr = arr
        self.s.nh = len(specs)
        self.s.hf = arr

    def hpower(self, k):
        p = np.zeros(self.ns)
        lib().orc_diag_hpower(C.byref(self.s), k, _ptr(p))
        return p

    def step(self, st, gamma0):
        if self.one4one and (st.marker & 4):
            # Gencore::run order: beam, sort, field, slippage (Gencore.cpp:114-147)
            lib().orc_beam_track(C.byref(self.s), C.byref(st), gamma0, st.dz)
            self.sort()
            lib().orc_field_advance(C.byref(self.s), C.byref(st), st.dz)
            lib().orc_slippage(C.byref(self.s), st.slip)
            return
        lib().orc_step_all(C.byref(self.s), C.byref(st), gamma0)

    def sort(self):
        """Beam::sort of a one4one beam (orc_sort_one4one), in place."""
        slen = 4.0 * np.arcsin(1.0) * self.s.slicelength / self.s.reflength
        n = lib().orc_sort_one4one(self.ns, self.off.ctypes.data_as(_lp), len(self.gamma), slen,
                                   self.s.slicelength, *[_ptr(getattr(self, k)) for k in (
                                       "gamma", "theta", "x", "y", "px", "py", "current")])
        assert n >= 0
        cnt = np.diff(self.off)
        if len(cnt) and cnt.max() > len(self.sc_ez):
            self.sc_ez = np.zeros(int(cnt.max()))
            self.s.sc_ez = _ptr(self.sc_ez)
        return n

    def run(self, steps, gamma0, record=("power", "bunching", "energy"), nsteps=None):
        nsteps = len(steps) if nsteps is None else nsteps
        rec = {k: [] for k in record}
        d = self.diag()
        for k in record:
            rec[k].append(d[k].copy() if k != "global_energy" else d[k])
        for i in range(nsteps):
            if steps[i].marker & 8:
                break
            self.step(steps[i], gamma0)
            d = self.diag()
            for k in record:
                rec[k].append(d[k].copy() if k != "global_energy" else d[k])
        return {k: np.array(v) for k, v in rec.items()}


def load_beam(sim, slice_params, npart_per_slice, nbins, shotnoise_seed=None, islice0=0,
              lambda0=None, sample=1, one_rank=0):
    """Load every slice as LoadBeam::init (LoadBeam.cpp:169-212): quiet
    loading, then shot noise from one sequential ran2 stream for
    time-dependent runs.  slice_params: list of dicts (per slice) or one
    dict; npart_per_slice: int or per-slice list."""
    L = lib()
    ns = sim.ns
    theta0 = 4.0 * np.arcsin(1.0)
    sn = L.orc_shotnoise_new(shotnoise_seed, one_rank) if shotnoise_seed is not None else None
    try:
        for j in range(ns):
            p = slice_params[j] if isinstance(slice_params, (list, tuple)) else slice_params
            sp = SliceParams(**{k: float(p.get(k, 0.0)) for k in (
                "gamma", "delgam", "current", "ex", "ey", "xcen", "ycen", "pxcen", "pycen",
                "betax", "betay", "alphax", "alphay", "bunch", "bunchphase", "emod",
                "emodphase")})
            a, b = int(sim.off[j]), int(sim.off[j + 1])
            n = b - a
            arrs = [getattr(sim, k)[a:b] for k in ("gamma", "theta", "x", "y", "px", "py")]
            L.orc_quiet_load(*[_ptr(v) for v in arrs], C.byref(sp), n, nbins, theta0,
                             j + islice0)
            ne = sp.current * lambda0 * sample / CE
            if sn is not None:
                L.orc_shotnoise(sn, _ptr(arrs[1]), n, nbins, ne)
            sim.current[j] = sp.current
    finally:
        if sn is not None:
            L.orc_shotnoise_free(sn)


def one4one_counts(currents, lambda0, sample=1):
    """Electrons per slice, round(I lambda0 sample / ce) (LoadBeam.cpp:172, 191)."""
    return [int(round(c * lambda0 * sample / CE)) for c in currents]


def load_beam_one4one(sim, slice_params, seed, lambda0, sample=1, rank=0):
    """one4one loading of every slice (orc_quiet_load_one4one); the Sim's
    counts must be one4one_counts of the currents."""
    ns = sim.ns
    sps = (SliceParams * ns)()
    for j in range(ns):
        p = slice_params[j] if isinstance(slice_params, (list, tuple)) else slice_params
        sps[j] = SliceParams(**{k: float(p.get(k, 0.0)) for k in (
            "gamma", "delgam", "current", "ex", "ey", "xcen", "ycen", "pxcen", "pycen",
            "betax", "betay", "alphax", "alphay", "bunch", "bunchphase", "emod", "emodphase")})
        sim.current[j] = sps[j].current
    theta0 = 4.0 * np.arcsin(1.0) * sample
    lib().orc_quiet_load_one4one(seed, rank, ns, sim.off.ctypes.data_as(_lp), sps, theta0,
                                 *[_ptr(getattr(sim, k)) for k in ("gamma", "theta", "x", "y", "px", "py")])
    sim.one4one = True
    sim.s.nbins = 1


def load_gauss_all(sim, power, w0, lambda0, z0=0.0):
    gm = sim.s.gridmax
    for j in range(sim.ns):
        f = np.zeros(sim.s.ngrid * sim.s.ngrid * 2)
        lib().orc_load_gauss(_ptr(f), sim.s.ngrid, gm, lambda0, power, w0, z0, 0.0, 0.0, 0.0)
        sim.field[j] = f.reshape(-1, 2)


# ---------------------------------------------------------------------------
# &alter_setup conversions (src/Main/AlterSetup.cpp:136-209), restated on
# numpy arrays: fields are [nslice][cells][2] in storage order with ring
# offset `first`, beams are CSR (off[nslice+1]) with coordinate arrays.
# ---------------------------------------------------------------------------
def field_harmonic_conversion(field, first, h, resample):
    """Field::harmonicConversion (src/Core/Field.cpp:261-288): each storage
    slot copied into h slots, first *= h.  Returns (field, first)."""
    if not resample or h == 1:
        return field.copy(), first
    nloc = field.shape[0]
    out = np.zeros((nloc * h,) + field.shape[1:])
    for i in range(nloc, 0, -1):
        for j in range(h):
            out[i * h - 1 - j] = field[i - 1]
    return out, first * h


def field_subharmonic_conversion(field, first, h, resample):
    """Field::subharmonicConversion (src/Core/Field.cpp:188-258): groups of h
    slices in ring order summed into the group's first slot, compacted to
    slot i from i h + first % h, and scaled to the group's mean power.
    Returns (field, first)."""
    if not resample or h == 1:
        return field.copy(), first
    f = field.copy()
    nsize = f.shape[0]
    pw = [float(np.sum(f[i] * f[i])) for i in range(nsize)]
    for i in range(0, nsize, h):
        is0 = (i + first) % nsize
        for j in range(1, h):
            is1 = (i + j + first) % nsize
            f[is0] += f[is1]
            pw[is0] += pw[is1]
    scl = 1.0 / h
    di = first % h
    for i in range(nsize // h):
        f[i] = f[i * h + di]
        pw[i] = pw[i * h + di] * scl
    first = (first - di) // h
    f = f[:nsize // h].copy()
    for i in range(f.shape[0]):
        pl = float(np.sum(f[i] * f[i]))
        if pw[i] > 0:
            f[i] *= np.sqrt(pw[i] / pl)
    return f, first


def beam_harmonic_conversion(off, theta, h, resample):
    """Beam::harmonicConversion (src/Core/Beam.cpp:187-241) before its sort:
    theta *= h; resampling moves slice i to slice h i.  Returns (off, theta)."""
    theta = theta * float(h)
    if not resample:
        return off.copy(), theta
    ns = len(off) - 1
    cnt = np.zeros(ns * h, np.int64)
    cnt[::h] = np.diff(off)
    new = np.zeros(ns * h + 1, np.int64)
    new[1:] = np.cumsum(cnt)
    return new, theta


def beam_subharmonic_conversion(off, theta, h, resample, lambda0, sample):
    """Beam::subharmonicConversion (src/Core/Beam.cpp:243-299) before its
    sort: theta /= h; resampling puts every particle into slice 0 at
    theta + i dtheta (dtheta the old slice in the new phase units) and
    keeps ns / h slices.  lambda0 and sample are the values before the
    conversion.  Returns (off, theta)."""
    theta = theta / float(h)
    if not resample:
        return off.copy(), theta
    ns = len(off) - 1
    assert ns % h == 0
    slicelength, reflength = lambda0 * float(sample) * h, lambda0 * h
    dtheta = 4.0 * np.arcsin(1.0) * slicelength / reflength / float(h)
    for i in range(1, ns):
        theta[off[i]:off[i + 1]] += i * dtheta
    new = np.full(ns // h + 1, off[-1], np.int64)
    new[0] = 0
    return new, theta
