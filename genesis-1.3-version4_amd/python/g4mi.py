"""ctypes binding of libg4mi.so, the MI355X undulator-step library.

This module only marshals arrays across the C-ABI declared in
include/g4mi.h; every computation happens in the library (HIP kernels on
the GPU, host input generation in C++).  Loading fails loudly if the
library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
# G4MI_LIBRARY names another build of the same library (e.g. a debug build)
LIB_PATH = os.environ.get("G4MI_LIBRARY") or os.path.join(PKG, "lib", "libg4mi.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "g4mi.h")

FP64, FP32 = 0, 1
CE = 4.8032045e-11

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_longlong)
_ip = C.POINTER(C.c_int)

STEP_FIELDS = ("aw", "ax", "ay", "ku", "kx", "ky", "gradx", "grady", "qf", "qx", "qy", "cx",
               "cy", "chic_angle", "chic_lb", "chic_ld", "chic_lt", "slip", "phaseshift", "z",
               "dz")


class StepParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in STEP_FIELDS] + [("helical", C.c_int),
                                                          ("marker", C.c_int)]


SLICE_FIELDS = ("gamma", "delgam", "current", "ex", "ey", "xcen", "ycen", "pxcen", "pycen",
                "betax", "betay", "alphax", "alphay", "bunch", "bunchphase", "emod", "emodphase")


class SliceBeam(C.Structure):
    _fields_ = [(n, C.c_double) for n in SLICE_FIELDS]


DIAG_FIELDS = ("power", "bunching", "bunchingphase", "energy", "energyspread", "xposition",
               "yposition", "xsize", "ysize", "f_xposition", "f_yposition", "f_xsize",
               "f_ysize", "nf_intensity", "ff_intensity", "global_energy",
               "pxposition", "pyposition", "nf_phase", "ff_phase", "efield", "wakefield",
               "lscfield", "sscfield", "xmin", "xmax", "pxmin", "pxmax", "ymin", "ymax",
               "pymin", "pymax", "emin", "emax",
               "g_energy", "g_energyspread", "g_xposition", "g_xsize", "g_yposition", "g_ysize",
               "g_f_xposition", "g_f_xsize", "g_f_yposition", "g_f_ysize", "g_nf_intensity",
               "g_ff_intensity",
               "xdivergence", "ydivergence", "xpointing", "ypointing",
               "g_xdivergence", "g_ydivergence", "g_xpointing", "g_ypointing")
# outputs of one value per record rather than one per slice
DIAG_SCALARS = ("global_energy", "g_energy", "g_energyspread", "g_xposition", "g_xsize",
                "g_yposition", "g_ysize", "g_f_xposition", "g_f_xsize", "g_f_yposition",
                "g_f_ysize", "g_nf_intensity", "g_ff_intensity",
                "g_xdivergence", "g_ydivergence", "g_xpointing", "g_ypointing")


SPECTRAL_FIELDS = ("xdivergence", "ydivergence", "xpointing", "ypointing",
                   "g_xdivergence", "g_ydivergence", "g_xpointing", "g_ypointing")


class DiagOut(C.Structure):
    _fields_ = [(n, _dp) for n in DIAG_FIELDS]


class G4miError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise G4miError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.g4mi_last_error.restype = C.c_char_p
        L.g4mi_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, _lp, C.c_int,
                                  C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
        L.g4mi_destroy.argtypes = [C.c_void_p]
        L.g4mi_set_efield.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.g4mi_set_incoherent.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.g4mi_incoherent_state.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.c_int]
        L.g4mi_diag_bunching.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.g4mi_set_sorting.argtypes = [C.c_void_p, C.c_int]
        L.g4mi_add_harmonic.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int)]
        L.g4mi_harmonic_upload.argtypes = [C.c_void_p, C.c_int, _dp]
        L.g4mi_harmonic_download.argtypes = [C.c_void_p, C.c_int, _dp]
        L.g4mi_harmonic_power.argtypes = [C.c_void_p, C.c_int, _dp]
        L.g4mi_sort.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
        L.g4mi_get_slices.argtypes = [C.c_void_p, _lp, _dp]
        L.g4mi_one4one_counts.argtypes = [C.c_int, C.c_void_p, C.c_double, C.c_double, _lp]
        L.g4mi_load_beam_one4one.argtypes = [C.c_int, _lp, C.c_void_p, C.c_int, C.c_double] + [_dp] * 6
        L.g4mi_set_wake.argtypes = [C.c_void_p, _dp]
        L.g4mi_set_field_solver.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                            C.c_double]
        L.g4mi_beam_upload.argtypes = [C.c_void_p] + [_dp] * 7
        L.g4mi_beam_download.argtypes = [C.c_void_p] + [_dp] * 6
        L.g4mi_field_upload.argtypes = [C.c_void_p, _dp, C.c_int]
        L.g4mi_field_download.argtypes = [C.c_void_p, _dp, _ip]
        L.g4mi_beam_track.argtypes = [C.c_void_p, C.POINTER(StepParams), C.c_double, C.c_double]
        L.g4mi_field_track.argtypes = [C.c_void_p, C.POINTER(StepParams), C.c_double]
        L.g4mi_slippage.argtypes = [C.c_void_p, C.c_double]
        L.g4mi_step.argtypes = [C.c_void_p, C.POINTER(StepParams), C.c_double]
        L.g4mi_run.argtypes = [C.c_void_p, C.POINTER(StepParams), C.c_int, C.c_double,
                               C.POINTER(DiagOut)]
        L.g4mi_diag.argtypes = [C.c_void_p, C.POINTER(DiagOut)]
        L.g4mi_sort_index.argtypes = [C.c_void_p, C.c_double, _ip, _lp]
        L.g4mi_sort_index_noalloc.argtypes = [C.c_void_p, C.c_double, _ip, _lp]
        L.g4mi_sync.argtypes = [C.c_void_p]
        L.g4mi_set_timing.argtypes = [C.c_void_p, C.c_int]
        L.g4mi_set_schedule.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.g4mi_get_schedule.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.g4mi_kernel_times.argtypes = [C.c_void_p, _dp, C.c_int, _lp]
        L.g4mi_parse_lattice.argtypes = [C.c_char_p, C.c_char_p, C.c_double, C.c_double,
                                         C.c_double, C.POINTER(StepParams), C.c_int, _ip]
        L.g4mi_load_beam.argtypes = [C.c_int, _lp, C.POINTER(SliceBeam), C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_double, C.c_double, C.c_int] + [_dp] * 6
        L.g4mi_load_gauss.argtypes = [_dp, C.c_int] + [C.c_double] * 6
        L.g4mi_load_beam_device.argtypes = [C.c_void_p, C.POINTER(SliceBeam), C.c_int, C.c_int, C.c_int,
                                            C.c_int, C.c_double, C.c_double, C.c_int]
        L.g4mi_run_deck.argtypes = [C.c_char_p, C.c_int, C.c_int]
        _fp = C.POINTER(C.c_float)
        L.g4mi_fno_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_uint]
        L.g4mi_fno_destroy.argtypes = [C.c_void_p]
        L.g4mi_fno_param_count.argtypes = [C.c_void_p]
        L.g4mi_fno_param_count.restype = C.c_longlong
        L.g4mi_fno_get_params.argtypes = [C.c_void_p, _fp]
        L.g4mi_fno_set_params.argtypes = [C.c_void_p, _fp]
        L.g4mi_fno_forward.argtypes = [C.c_void_p, _fp, _fp, C.c_int]
        L.g4mi_fno_loss_grad.argtypes = [C.c_void_p, _fp, _fp, C.c_int, _dp, _fp]
        L.g4mi_fno_reserve.argtypes = [C.c_void_p, C.c_longlong]
        L.g4mi_fno_pair_count.argtypes = [C.c_void_p]
        L.g4mi_fno_pair_count.restype = C.c_longlong
        L.g4mi_fno_train.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_longlong, C.c_uint,
                                     _dp, _dp]
        L.g4mi_fno_eval.argtypes = [C.c_void_p, C.c_longlong, C.c_longlong, C.c_int, _dp]
        L.g4mi_attach_fno.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double]
        L.g4mi_fno_stats.argtypes = [C.c_void_p, _lp, _lp, _dp, _dp]
        L.g4mi_fno_train_span.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                          C.c_longlong, C.c_uint, _dp, _dp]
        L.g4mi_fno_judge.argtypes = [C.c_void_p, _lp, _dp, _dp, _dp]
        L.g4mi_run_deck.argtypes = [C.c_char_p, C.c_int, C.c_int]
        L.g4mi_check_deck.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.g4mi_wake_loss.argtypes = ([C.c_int, C.c_int, C.c_double, _dp, _dp, C.c_double, C.c_int]
                                     + [C.c_double] * 6 + [_dp] * 4)
        L.g4mi_h5_read.argtypes = [C.c_char_p, C.c_char_p, _dp, C.c_longlong, _lp]
        L.g4mi_h5_read.restype = C.c_longlong
        _lib = L
    return _lib


def _chk(rc, what):
    if rc < 0:
        raise G4miError(f"{what} failed ({rc}): {lib().g4mi_last_error().decode()}")
    return rc


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


# ---------------------------------------------------------------- input side
def parse_lattice(text, line, delz, lambda0, gamma0):
    n = C.c_int(0)
    _chk(lib().g4mi_parse_lattice(text.encode(), line.encode(), delz, lambda0, gamma0, None, 0,
                                  C.byref(n)), "g4mi_parse_lattice")
    out = (StepParams * max(1, n.value))()
    _chk(lib().g4mi_parse_lattice(text.encode(), line.encode(), delz, lambda0, gamma0, out,
                                  n.value, C.byref(n)), "g4mi_parse_lattice")
    return (StepParams * n.value)(*out[:n.value])


def load_beam(off, slices, nbins, shotnoise, seed, lambda0, sample=1.0, timerun=True,
              islice0=0):
    """Quiet-load every slice (and apply shot noise) as LoadBeam::init.
    slices: list of dicts with SLICE_FIELDS keys.  Returns dict of SoA."""
    off = np.ascontiguousarray(off, dtype=np.int64)
    ns = len(off) - 1
    sb = (SliceBeam * ns)(*[SliceBeam(**{k: float(s.get(k, 0.0)) for k in SLICE_FIELDS})
                            for s in slices])
    n = int(off[-1])
    arrs = {k: np.zeros(n) for k in ("gamma", "theta", "x", "y", "px", "py")}
    _chk(lib().g4mi_load_beam(ns, off.ctypes.data_as(_lp), sb, nbins, int(shotnoise), seed,
                              islice0, lambda0, sample, int(timerun),
                              *[_p(arrs[k]) for k in ("gamma", "theta", "x", "y", "px", "py")]),
         "g4mi_load_beam")
    return arrs


def one4one_counts(slices, lambda0, sample=1.0):
    """Electrons per slice of a one4one beam (g4mi_one4one_counts)."""
    ns = len(slices)
    sb = (SliceBeam * ns)(*[SliceBeam(**{k: float(s.get(k, 0.0)) for k in SLICE_FIELDS}) for s in slices])
    cnt = np.zeros(ns, dtype=np.int64)
    _chk(lib().g4mi_one4one_counts(ns, sb, lambda0, sample, cnt.ctypes.data_as(_lp)), "g4mi_one4one_counts")
    return cnt


def load_beam_one4one(off, slices, seed, sample=1.0):
    """one4one loading of every slice from one ran2 stream (g4mi_load_beam_one4one)."""
    off = np.ascontiguousarray(off, dtype=np.int64)
    ns = len(off) - 1
    sb = (SliceBeam * ns)(*[SliceBeam(**{k: float(s.get(k, 0.0)) for k in SLICE_FIELDS}) for s in slices])
    n = int(off[-1])
    arrs = {k: np.zeros(n) for k in ("gamma", "theta", "x", "y", "px", "py")}
    _chk(lib().g4mi_load_beam_one4one(ns, off.ctypes.data_as(_lp), sb, seed, sample,
                                      *[_p(arrs[k]) for k in ("gamma", "theta", "x", "y", "px", "py")]),
         "g4mi_load_beam_one4one")
    return arrs


def load_gauss(ngrid, gridmax, lambda0, power, waist, waist_pos=0.0, phase=0.0):
    f = np.zeros(ngrid * ngrid * 2)
    _chk(lib().g4mi_load_gauss(_p(f), ngrid, gridmax, lambda0, power, waist, waist_pos, phase),
         "g4mi_load_gauss")
    return f.reshape(ngrid * ngrid, 2)


# ---------------------------------------------------------------- device side
class Context:
    """A g4mi device context: one time window of slices on one GPU."""

    def __init__(self, off, ngrid, gridmax, lambda0, sample=1.0, timerun=True, periodic=False,
                 precision=FP64, device=0):
        self.off = np.ascontiguousarray(off, dtype=np.int64)
        self.ns = len(self.off) - 1
        self.ngrid = ngrid
        self.n = int(self.off[-1])
        self.precision = precision
        h = C.c_void_p()
        _chk(lib().g4mi_create(C.byref(h), device, self.ns, self.off.ctypes.data_as(_lp), ngrid,
                               gridmax, lambda0, sample, int(timerun), int(periodic), precision),
             "g4mi_create")
        self.h = h

    def close(self):
        if self.h:
            lib().g4mi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_efield(self, longrange=False, nz=0, nphi=0, ngrid_r=100, rmax=0.0):
        _chk(lib().g4mi_set_efield(self.h, int(longrange), nz, nphi, ngrid_r, rmax),
             "g4mi_set_efield")

    def set_incoherent(self, seed=1234, nbins=4, do_loss=False, do_spread=False):
        """Beam::initIncoherent (src/Core/Incoherent.cpp:17-31), rank 0."""
        _chk(lib().g4mi_set_incoherent(self.h, seed, nbins, int(do_loss), int(do_spread)),
             "g4mi_set_incoherent")

    def incoherent_state(self, state=None):
        """Read (state None) or set the incoherent stream's 35-int state."""
        buf = (C.c_int * 35)()
        if state is not None:
            buf[:] = [int(v) for v in state]
        _chk(lib().g4mi_incoherent_state(self.h, buf, 0 if state is None else 1), "g4mi_incoherent_state")
        return list(buf)

    def add_harmonic(self, harm, ngrid, gridmax):
        """A field at harmonic `harm` with its own grid; returns its index."""
        idx = C.c_int(-1)
        _chk(lib().g4mi_add_harmonic(self.h, harm, ngrid, gridmax, C.byref(idx)), "g4mi_add_harmonic")
        self._hgrid = getattr(self, "_hgrid", {})
        self._hgrid[idx.value] = ngrid
        return idx.value

    def upload_harmonic(self, index, field):
        f = np.ascontiguousarray(field, dtype=np.float64)
        assert f.size == self.ns * self._hgrid[index] ** 2 * 2
        _chk(lib().g4mi_harmonic_upload(self.h, index, _p(f)), "g4mi_harmonic_upload")

    def download_harmonic(self, index):
        ng = self._hgrid[index]
        f = np.zeros((self.ns, ng * ng, 2))
        _chk(lib().g4mi_harmonic_download(self.h, index, _p(f)), "g4mi_harmonic_download")
        return f

    def harmonic_power(self, index):
        p = np.zeros(self.ns)
        _chk(lib().g4mi_harmonic_power(self.h, index, _p(p)), "g4mi_harmonic_power")
        return p

    def set_sorting(self, one4one=True):
        """Beam::initSorting: Beam::sort at sort markers for one4one beams."""
        _chk(lib().g4mi_set_sorting(self.h, int(one4one)), "g4mi_set_sorting")

    def sort(self):
        """Beam::sort now; returns the particle count."""
        n = C.c_longlong(0)
        _chk(lib().g4mi_sort(self.h, C.byref(n)), "g4mi_sort")
        return n.value

    def slices(self):
        """Current slice offsets (ns+1) and currents (ns)."""
        off = np.zeros(self.ns + 1, dtype=np.int64)
        cur = np.zeros(self.ns)
        _chk(lib().g4mi_get_slices(self.h, off.ctypes.data_as(_lp), _p(cur)), "g4mi_get_slices")
        return off, cur

    def set_wake(self, eloss):
        """Per-slice wake loss in eV/m (Collective::apply), or None."""
        if eloss is None:
            _chk(lib().g4mi_set_wake(self.h, None), "g4mi_set_wake")
            return
        e = np.ascontiguousarray(eloss, dtype=np.float64)
        if e.shape != (self.ns,):
            raise ValueError("eloss needs one value per slice")
        _chk(lib().g4mi_set_wake(self.h, _p(e)), "g4mi_set_wake")

    def set_field_solver(self, fft=False, source_filter=False, xcut=1.0, ycut=1.0, sigmoid=1.0):
        """Field::initSolver (src/Core/Field.cpp:72-83): the ADI (fft=False)
        or the spectral solver, with FieldSolverFFT::initSourceFilter's
        settings."""
        _chk(lib().g4mi_set_field_solver(self.h, int(fft), int(source_filter), xcut, ycut, sigmoid),
             "g4mi_set_field_solver")

    def upload_beam(self, beam, current):
        cur = np.ascontiguousarray(current, dtype=np.float64)
        a = [np.ascontiguousarray(beam[k], dtype=np.float64)
             for k in ("gamma", "theta", "x", "y", "px", "py")]
        _chk(lib().g4mi_beam_upload(self.h, *[_p(v) for v in a], _p(cur)), "g4mi_beam_upload")

    def load_beam(self, slices, nbins, shotnoise, seed, lambda0, sample=1.0, timerun=True, islice0=0):
        """LoadBeam::init into the device arrays (g4mi_load_beam_device): the
        beam of load_beam() with the context's slice counts, uploaded with
        the slices' currents."""
        if len(slices) != self.ns:
            raise ValueError(f"load_beam: {len(slices)} slice records for {self.ns} slices")
        sb = (SliceBeam * self.ns)(*[SliceBeam(**{k: float(s.get(k, 0.0)) for k in SLICE_FIELDS})
                                     for s in slices])
        _chk(lib().g4mi_load_beam_device(self.h, sb, nbins, int(shotnoise), seed, islice0, lambda0, sample,
                                         int(timerun)), "g4mi_load_beam_device")

    def download_beam(self):
        """The particles in their current slice order (after one4one sorting
        the counts per slice are those of slices())."""
        off = np.zeros(self.ns + 1, dtype=np.int64)
        _chk(lib().g4mi_get_slices(self.h, off.ctypes.data_as(_lp), None), "g4mi_get_slices")
        out = {k: np.zeros(int(off[-1])) for k in ("gamma", "theta", "x", "y", "px", "py")}
        _chk(lib().g4mi_beam_download(self.h, *[_p(out[k]) for k in
                                                ("gamma", "theta", "x", "y", "px", "py")]),
             "g4mi_beam_download")
        return out

    def upload_field(self, field, first=0):
        f = np.ascontiguousarray(field, dtype=np.float64)
        assert f.size == self.ns * self.ngrid * self.ngrid * 2
        _chk(lib().g4mi_field_upload(self.h, _p(f), first), "g4mi_field_upload")

    def download_field(self):
        f = np.zeros((self.ns, self.ngrid * self.ngrid, 2))
        first = C.c_int(0)
        _chk(lib().g4mi_field_download(self.h, _p(f), C.byref(first)), "g4mi_field_download")
        return f, first.value

    def beam_track(self, st, gamma0, delz):
        _chk(lib().g4mi_beam_track(self.h, C.byref(st), gamma0, delz), "g4mi_beam_track")

    def field_track(self, st, delz):
        _chk(lib().g4mi_field_track(self.h, C.byref(st), delz), "g4mi_field_track")

    def slippage(self, slip):
        _chk(lib().g4mi_slippage(self.h, slip), "g4mi_slippage")

    def step(self, st, gamma0):
        _chk(lib().g4mi_step(self.h, C.byref(st), gamma0), "g4mi_step")

    def sync(self):
        _chk(lib().g4mi_sync(self.h), "g4mi_sync")

    def diag_bunching(self, hmax):
        """Bunching and its phase at harmonics 2..hmax, arrays [hmax-1][ns]."""
        b = np.zeros((hmax - 1, self.ns))
        ph = np.zeros((hmax - 1, self.ns))
        _chk(lib().g4mi_diag_bunching(self.h, hmax, _p(b), _p(ph)), "g4mi_diag_bunching")
        return b, ph

    def diag(self, spectral=False):
        """One diagnostics record; spectral=True adds the divergence and
        pointing block (a 2-D DFT per slice)."""
        keys = [k for k in DIAG_FIELDS if spectral or k not in SPECTRAL_FIELDS]
        out = {k: np.zeros(1 if k in DIAG_SCALARS else self.ns) for k in keys}
        d = DiagOut(*[_p(out.get(k)) for k in DIAG_FIELDS])
        _chk(lib().g4mi_diag(self.h, C.byref(d)), "g4mi_diag")
        for k in DIAG_SCALARS:
            if k in out:
                out[k] = out[k][0]
        return out

    def run(self, steps, gamma0, record=("power", "bunching", "energy")):
        """Initial diagnostics, then every step with a record after each."""
        nst = len(steps)
        recs = {k: np.zeros((nst + 1, self.ns)) for k in record}
        d0 = self.diag()
        for k in record:
            recs[k][0] = d0[k]
        rows = (DiagOut * max(1, nst))()
        keep = []
        for i in range(nst):
            ptrs = []
            for k in DIAG_FIELDS:
                if k in recs:
                    ptrs.append(recs[k][i + 1].ctypes.data_as(_dp))
                else:
                    ptrs.append(None)
            rows[i] = DiagOut(*ptrs)
        arr = (StepParams * max(1, nst))(*steps) if nst else (StepParams * 1)()
        done = _chk(lib().g4mi_run(self.h, arr, nst, gamma0, rows), "g4mi_run")
        keep.append(rows)
        return {k: v[:done + 1] for k, v in recs.items()}

    def sort_index(self, slen):
        atar = np.zeros(self.n, dtype=np.int32)
        moved = C.c_longlong(0)
        _chk(lib().g4mi_sort_index(self.h, slen, atar.ctypes.data_as(_ip), C.byref(moved)),
             "g4mi_sort_index")
        return atar, moved.value

    def sort_index_noalloc(self, slen):
        """g4mi_sort_index_noalloc: the same targets from the library's static
        device scratch (at most 4194304 particles; no allocation)."""
        atar = np.zeros(self.n, dtype=np.int32)
        moved = C.c_longlong(0)
        _chk(lib().g4mi_sort_index_noalloc(self.h, slen, atar.ctypes.data_as(_ip), C.byref(moved)),
             "g4mi_sort_index_noalloc")
        return atar, moved.value

    def set_schedule(self, name, value):
        _chk(lib().g4mi_set_schedule(self.h, name.encode(), int(value)), "g4mi_set_schedule")

    def schedule(self):
        import json
        buf = C.create_string_buffer(4096)
        _chk(lib().g4mi_get_schedule(self.h, buf, 4096), "g4mi_get_schedule")
        return json.loads(buf.value.decode())

    def attach_fno(self, fno, mode, every=8, threshold=1e-3):
        """mode 0 detach, 1 capture training pairs, 2 replace the ADI (with fallback)."""
        _chk(lib().g4mi_attach_fno(self.h, fno.h if fno is not None else None, mode, every,
                                   threshold), "g4mi_attach_fno")

    def fno_judge(self):
        n = C.c_longlong()
        ef, eu, ra = C.c_double(), C.c_double(), C.c_double()
        _chk(lib().g4mi_fno_judge(self.h, C.byref(n), C.byref(ef), C.byref(eu), C.byref(ra)),
             "g4mi_fno_judge")
        return dict(judge_steps=n.value, rel_field_err=ef.value, rel_update_err=eu.value,
                    mean_residual_all=ra.value)

    def fno_stats(self):
        a, b = C.c_longlong(), C.c_longlong()
        r1, r2 = C.c_double(), C.c_double()
        _chk(lib().g4mi_fno_stats(self.h, C.byref(a), C.byref(b), C.byref(r1), C.byref(r2)),
             "g4mi_fno_stats")
        return dict(fno_steps=a.value, adi_steps=b.value, mean_residual_fno=r1.value,
                    mean_residual_adi=r2.value)

    TIMER_NAMES = ("transverse", "space_charge", "push", "deposit", "adi", "slip_diag")

    def set_timing(self, on=True):
        _chk(lib().g4mi_set_timing(self.h, int(on)), "g4mi_set_timing")

    def kernel_times(self):
        """Average device ms per step of each phase since set_timing(True)."""
        ms = np.zeros(len(self.TIMER_NAMES))
        n = C.c_longlong()
        _chk(lib().g4mi_kernel_times(self.h, _p(ms), len(ms), C.byref(n)), "g4mi_kernel_times")
        out = dict(zip(self.TIMER_NAMES, ms.tolist()))
        out["steps"] = n.value
        return out


class Fno:
    """The learned field step (a Fourier neural operator on one slice's grid)."""

    def __init__(self, ngrid, width=32, modes=16, layers=4, hidden=64, max_batch=16, seed=1,
                 device=0):
        self.ngrid, self.width, self.modes = ngrid, width, modes
        self.layers, self.hidden, self.max_batch = layers, hidden, max_batch
        h = C.c_void_p()
        _chk(lib().g4mi_fno_create(C.byref(h), device, ngrid, width, modes, layers, hidden,
                                   max_batch, seed), "g4mi_fno_create")
        self.h = h

    def close(self):
        if self.h:
            lib().g4mi_fno_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def param_count(self):
        return int(lib().g4mi_fno_param_count(self.h))

    def get_params(self):
        p = np.zeros(self.param_count(), dtype=np.float32)
        _chk(lib().g4mi_fno_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.param_count()
        _chk(lib().g4mi_fno_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))), "set_params")

    def forward(self, inp):
        inp = np.ascontiguousarray(inp, dtype=np.float32)
        B = inp.shape[0]
        out = np.zeros((B, self.ngrid * self.ngrid, 2), dtype=np.float32)
        _chk(lib().g4mi_fno_forward(self.h, inp.ctypes.data_as(C.POINTER(C.c_float)),
                                    out.ctypes.data_as(C.POINTER(C.c_float)), B), "fno_forward")
        return out

    def loss_grad(self, inp, tgt):
        inp = np.ascontiguousarray(inp, dtype=np.float32)
        tgt = np.ascontiguousarray(tgt, dtype=np.float32)
        loss = C.c_double()
        g = np.zeros(self.param_count(), dtype=np.float32)
        _chk(lib().g4mi_fno_loss_grad(self.h, inp.ctypes.data_as(C.POINTER(C.c_float)),
                                      tgt.ctypes.data_as(C.POINTER(C.c_float)), inp.shape[0],
                                      C.byref(loss), g.ctypes.data_as(C.POINTER(C.c_float))),
             "fno_loss_grad")
        return loss.value, g

    def reserve(self, n):
        _chk(lib().g4mi_fno_reserve(self.h, n), "fno_reserve")

    def pair_count(self):
        return int(lib().g4mi_fno_pair_count(self.h))

    def train(self, epochs, batch, lr, nholdout=0, seed=1, first_epoch=0, total_epochs=None):
        """Adam over `epochs` epochs; with first_epoch/total_epochs, those
        epochs of a cosine schedule spanning total_epochs."""
        tl = np.zeros(max(1, epochs))
        hl = np.zeros(max(1, epochs))
        if total_epochs is None and first_epoch == 0:
            _chk(lib().g4mi_fno_train(self.h, epochs, batch, lr, nholdout, seed, _p(tl), _p(hl)),
                 "fno_train")
        else:
            total = total_epochs if total_epochs is not None else first_epoch + epochs
            _chk(lib().g4mi_fno_train_span(self.h, first_epoch, epochs, total, batch, lr, nholdout,
                                           seed, _p(tl), _p(hl)), "fno_train_span")
        return tl[:epochs], hl[:epochs]

    def eval(self, first, last, batch):
        loss = C.c_double()
        _chk(lib().g4mi_fno_eval(self.h, first, last, batch, C.byref(loss)), "fno_eval")
        return loss.value

    def param_blocks(self):
        """(name, offset, size) of every parameter block, in storage order."""
        C_, H, nm2 = self.width, self.hidden, self.modes * self.modes
        out, o = [], 0
        for name, n in [("P", 4 * C_), ("bP", C_)]:
            out.append((name, o, n))
            o += n
        for l in range(self.layers):
            for name, n in [(f"Rr{l}", nm2 * C_ * C_), (f"Ri{l}", nm2 * C_ * C_),
                            (f"W{l}", C_ * C_), (f"b{l}", C_)]:
                out.append((name, o, n))
                o += n
        for name, n in [("Q1", C_ * H), ("bQ1", H), ("Q2", 2 * H), ("bQ2", 2)]:
            out.append((name, o, n))
            o += n
        return out


def run_deck(path, device=0, quiet=True):
    """Run a GENESIS 1.3 v4 input deck on the device (writes <rootname>.out.h5)."""
    _chk(lib().g4mi_run_deck(str(path).encode(), device, int(quiet)), "g4mi_run_deck")


def wake_loss(current, ds, sample=1, loss=None, radius=2.5e-3, roundpipe=True, conductivity=0.0,
              relaxation=0.0, gap=0.0, lgap=1.0, hrough=0.0, lrough=1.0):
    """Per-slice wake loss [eV/m] for a current profile (A per slice) and
    the single-particle wakes (resistive, geometric, roughness) on the
    len(current) * sample bins of length ds (g4mi_wake_loss)."""
    cur = np.ascontiguousarray(current, dtype=np.float64)
    ns = len(cur)
    ext = None if loss is None else np.ascontiguousarray(np.broadcast_to(loss, (ns,)), dtype=np.float64)
    eloss = np.zeros(ns)
    w = [np.zeros(ns * sample) for _ in range(3)]
    _chk(lib().g4mi_wake_loss(ns, sample, ds, _p(cur), _p(ext), radius, int(roundpipe), conductivity,
                              relaxation, gap, lgap, hrough, lrough, _p(eloss), *[_p(a) for a in w]),
         "g4mi_wake_loss")
    return eloss, {"resistive": w[0], "geometric": w[1], "roughness": w[2]}


def check_deck(path):
    """Parse and set up a deck without a device; summary of its last &track."""
    import json
    buf = C.create_string_buffer(4096)
    _chk(lib().g4mi_check_deck(str(path).encode(), buf, 4096), "g4mi_check_deck")
    return json.loads(buf.value.decode()) if buf.value else {}


def h5_read(file, dataset):
    """A double dataset of an HDF5 file as a numpy array (1-D or 2-D)."""
    dims = np.zeros(2, dtype=np.int64)
    n = lib().g4mi_h5_read(str(file).encode(), dataset.encode(), None, 0,
                           dims.ctypes.data_as(_lp))
    _chk(n, "g4mi_h5_read")
    out = np.zeros(int(n))
    _chk(lib().g4mi_h5_read(str(file).encode(), dataset.encode(), _p(out), int(n),
                            dims.ctypes.data_as(_lp)), "g4mi_h5_read")
    return out.reshape(int(dims[0]), int(dims[1])) if dims[1] > 0 else out


def steps_from(seq):
    """Convert any sequence of ctypes step records with the same field
    names (e.g. the oracle's) to StepParams."""
    out = (StepParams * len(seq))()
    for i, s in enumerate(seq):
        for k in STEP_FIELDS + ("helical", "marker"):
            setattr(out[i], k, getattr(s, k))
    return out
