/*
 * g4mi -- MI355X-native (gfx950) undulator integration step for GENESIS 1.3 v4.
 *
 * C-ABI of the drop-in boundary.  Plain pointers and sizes only; every entry
 * point returns 0 on success or a negative G4MI_E* code and never throws.
 * Device memory is owned by the context; host pointers are borrowed for the
 * duration of the call.  One host thread per context; all work is ordered on
 * the context's HIP stream; download/diag calls synchronise that stream.
 *
 * The reference has no C-ABI for this path.  Each entry point below names
 * the reference member function (file:line) whose contract it replaces; a
 * maintainer binds them from FieldSolver/BeamSolver subclasses whose
 * init/advance/track call the matching entry points (include/g4mi.hpp
 * wraps them for C++).
 */
#ifndef G4MI_H
#define G4MI_H

#ifdef __cplusplus
extern "C" {
#endif

#define G4MI_OK 0
#define G4MI_E_ARG -1      /* invalid argument or size                      */
#define G4MI_E_HIP -2      /* HIP runtime error                             */
#define G4MI_E_NODEV -3    /* no usable GPU                                 */
#define G4MI_E_STATE -4    /* call out of order (e.g. no beam uploaded)     */
#define G4MI_E_IO -5       /* input or output file error                    */
#define G4MI_E_PARSE -6    /* namelist or lattice syntax error              */

/* Arithmetic of a context. */
#define G4MI_FP64 0        /* plain FP64 form, follows the reference order  */
#define G4MI_FP32 1        /* FP32 push/field, FP64 slice base for theta,
                              FP32 ADI + one FP64 residual-correction sweep */

typedef struct g4mi_ctx g4mi_ctx;

/* One integration step of the unrolled lattice: the per-step arrays of the
 * reference's Undulator class (include/Undulator.h:73-77), one record per
 * step, as Lattice::generateLattice (src/Lattice/Lattice.cpp:53-149) fills
 * them. */
typedef struct {
    double aw, ax, ay, ku, kx, ky, gradx, grady;
    double qf, qx, qy, cx, cy;
    double chic_angle, chic_lb, chic_ld, chic_lt;
    double slip, phaseshift, z, dz;
    int helical, marker;
} g4mi_step_params;

/* Per-slice diagnostics of one output record (any pointer may be NULL).
 * Arrays have nslice entries in time-window order, as DiagBeam::getValues
 * (src/Core/Diagnostic.cpp:373-608) and DiagField::getValues (:718-996)
 * store them. */
typedef struct {
    double *power, *bunching, *bunchingphase, *energy, *energyspread;
    double *xposition, *yposition, *xsize, *ysize;
    double *f_xposition, *f_yposition, *f_xsize, *f_ysize;
    double *nf_intensity, *ff_intensity;
    double *global_energy; /* [1] */
    /* DiagBeam: mean momenta and the auxiliary block (LSC/SSC fields in
     * eV/m, wake field 0 in this build, coordinate extremes);
     * DiagField: near- and far-field phases */
    double *pxposition, *pyposition;
    double *nf_phase, *ff_phase;
    double *efield, *wakefield, *lscfield, *sscfield;
    double *xmin, *xmax, *pxmin, *pxmax, *ymin, *ymax, *pymin, *pymax, *emin, *emax;
    /* [1] each: current-weighted beam statistics over all slices
     * (Diagnostic.cpp:565-607) and power-weighted field statistics
     * (:906-969) */
    double *g_energy, *g_energyspread, *g_xposition, *g_xsize, *g_yposition, *g_ysize;
    double *g_f_xposition, *g_f_xsize, *g_f_yposition, *g_f_ysize;
    double *g_nf_intensity, *g_ff_intensity;
    /* DiagField's spectral block (Diagnostic.cpp:789-807, 839-846): rms
     * divergence and mean pointing angle per slice from |DFT(E)|^2, and
     * their power-weighted global values (:952-963) ([1] each).  Computed
     * only when one of these pointers is set (two DFT products per slice). */
    double *xdivergence, *ydivergence, *xpointing, *ypointing;
    double *g_xdivergence, *g_ydivergence, *g_xpointing, *g_ypointing;
} g4mi_diag_out;

/* Context on HIP device `device` holding a time window of `nslice` slices
 * with CSR particle offsets `off` (host, nslice+1 entries; uneven counts
 * allowed), a field of ngrid x ngrid cells with half width `gridmax`,
 * reference wavelength `lambda0` and slice length sample*lambda0.
 * Replaces Beam::init (src/Core/Beam.cpp:22-39) + Field::init
 * (src/Core/Field.cpp:34-70) + Control::init (src/Core/Control.cpp:108-154). */
int g4mi_create(g4mi_ctx **ctx, int device, int nslice, const long long *off, int ngrid,
                double gridmax, double lambda0, double sample, int timerun, int periodic,
                int precision);
int g4mi_destroy(g4mi_ctx *ctx);

/* Longitudinal space-charge settings, EFieldSolver::init
 * (src/Core/EFieldSolver.cpp:27-49) via Beam::initEField. */
int g4mi_set_efield(g4mi_ctx *ctx, int longrange, int nz, int nphi, int ngrid_r, double rmax);

/* Particle SoA and per-slice current (A), host -> device / device -> host.
 * Replaces the particle fill of LoadBeam::init (src/Loading/LoadBeam.cpp:202-211)
 * and the reads of Beam::beam by dumps. */
int g4mi_beam_upload(g4mi_ctx *ctx, const double *gamma, const double *theta, const double *x,
                     const double *y, const double *px, const double *py, const double *current);
int g4mi_beam_download(g4mi_ctx *ctx, double *gamma, double *theta, double *x, double *y,
                       double *px, double *py);

/* Field slices, interleaved complex [nslice][ngrid*ngrid][2] in storage
 * order, plus the ring offset `first` (src/Core/Field.h:57-59). */
int g4mi_field_upload(g4mi_ctx *ctx, const double *field, int first);
int g4mi_field_download(g4mi_ctx *ctx, double *field, int *first);

/* Beam::track (src/Core/Beam.cpp:97-112): half transverse step,
 * BeamSolver::advance (src/Core/BeamSolver.cpp:13-87) with long- and
 * short-range space charge, R56, half transverse step. */
int g4mi_beam_track(g4mi_ctx *ctx, const g4mi_step_params *st, double gammaref, double delz);

/* Field::track (src/Core/Field.cpp:177-185): FieldSolverADI::init
 * (src/Core/FieldSolverADI.cpp:125-182) when delz changes, then
 * FieldSolverADI::advance (:9-57): source deposition + ADI per slice; or,
 * after g4mi_set_field_solver(fft = 1), FieldSolverFFT::init/advance
 * (src/Core/FieldSolverFFT.cpp:9-60, 122-157): the same deposition, then
 * every slice propagated by exp(K2 delz) in Fourier space plus twice the
 * (optionally filtered) source. */
int g4mi_field_track(g4mi_ctx *ctx, const g4mi_step_params *st, double delz);

/* Incoherent synchrotron radiation, Beam::initIncoherent ->
 * Incoherent::init (src/Core/Incoherent.cpp:17-31) for rank 0, applied by
 * Incoherent::apply (:35-82) after the push of every undulator step: one
 * energy kick -dgamavg + dgamsig (2u - 1) per beamlet of `nbins`
 * particles, u from the single ran2 stream seeded with round(1e9 ran2(seed)),
 * drawn on the host in slice and particle order.  do_loss = do_spread = 0
 * switches it off.  FP64 form only. */
int g4mi_set_incoherent(g4mi_ctx *ctx, int seed, int nbins, int do_loss, int do_spread);
/* The state of the incoherent stream (35 ints: ran2's shuffle table, last
 * output and both generators): store = 0 reads it into state, store = 1
 * sets it.  No counterpart in the reference, whose Incoherent object lives
 * as long as the beam: it lets a host that rebuilds its context between
 * runs (a changed slice count or reference wavelength) continue the same
 * stream. */
int g4mi_incoherent_state(g4mi_ctx *ctx, int *state, int store);

/* Bunching at the harmonics 2..hmax of the beam's phase, DiagBeam::getValues
 * (src/Core/Diagnostic.cpp:431-436, 486-500: b_h = (1/N) sum exp(i h theta)
 * per slice, output bunching<h> = |b_h|, bunchingphase<h> = arg b_h).
 * bunching and phase are [hmax-1][nslice], row h-2 for harmonic h.  FP64
 * form; synchronous. */
int g4mi_diag_bunching(g4mi_ctx *ctx, int hmax, double *bunching, double *phase);

/* Fields at harmonics of the fundamental (&field harm = h: Field::init with
 * harm, src/Core/Field.cpp:34-70), FP64 form, ADI solver, at most 4 besides
 * the fundamental.  Each has its own grid (ngrid, gridmax) and wavenumber
 * h 2 pi / lambda0 and shares the fundamental's ring offset.  The push sums
 * their coupling fc(h) / xks_h awloc conj(E_h) exp(-i h theta) into the
 * ODE after the fundamental's (BeamSolver.cpp:23-32, 63-79, 144-163); each
 * odd harmonic gets the source of harmonic h (FieldSolverADI.cpp:19-50);
 * slippage zeroes their inflow slice too.  Fields are [nslice][ngrid^2]
 * complex in storage order (the fundamental's `first`); power is per beam
 * slice in W (DiagField). */
int g4mi_add_harmonic(g4mi_ctx *ctx, int harm, int ngrid, double gridmax, int *index);
int g4mi_harmonic_upload(g4mi_ctx *ctx, int index, const double *field);
int g4mi_harmonic_download(g4mi_ctx *ctx, int index, double *field);
int g4mi_harmonic_power(g4mi_ctx *ctx, int index, double *power);
/* DiagField's outputs of harmonic field `index` (power, sizes and
 * positions, intensities, far-field phase, their global values and
 * global_energy); the beam outputs of `out` are not touched. */
int g4mi_diag_harmonic(g4mi_ctx *ctx, int index, g4mi_diag_out *out);

/* one4one beams, Beam::initSorting (src/Core/Beam.cpp:75-79): with
 * one4one = 1 every step whose marker has bit 4 set runs Beam::sort
 * (:81-94) after the beam step, as Gencore::run (src/Core/Gencore.cpp:
 * 124-134): particles outside the time window [0, slen nslice] are dropped,
 * the others move to slice a + floor(theta / slen) with theta -= slen atar
 * (Sorting.cpp:74-145, 367-428), and each slice's current becomes
 * np ce / slicelength.  The order of the particles within a slice is the
 * source order (the reference's differs; the sets are the same).
 * g4mi_sort sorts now and returns the particle count; g4mi_get_slices
 * copies the current offsets (nslice+1) and currents (nslice).  FP64 form
 * only. */
int g4mi_set_sorting(g4mi_ctx *ctx, int one4one);
int g4mi_sort(g4mi_ctx *ctx, long long *nkept);
int g4mi_get_slices(g4mi_ctx *ctx, long long *off, double *current);

/* Wakefield energy loss, Collective::apply (src/Core/Collective.cpp:107-128):
 * eloss[nslice] in eV/m (Collective's wakeext + wakeint for each beam
 * slice), added as eloss * delz / eev to every particle's gamma after the
 * push of every step and reported as the `wakefield` diagnostic.  NULL
 * removes it.  FP64 form only. */
int g4mi_set_wake(g4mi_ctx *ctx, const double *eloss);

/* Wake potentials and the loss they give each slice (host): Wake::init's
 * single-particle wakes (src/Main/Wake.cpp:136-434: resistive wall of a
 * round or flat chamber with AC conductivity, geometric wake of periodic
 * gaps, surface roughness) on nslice*sample bins of length ds, convolved
 * with the current profile as Collective::initWake/update
 * (src/Core/Collective.cpp:36-196); eloss[nslice] = loss_ext (or 0) + the
 * convolution, in eV/m, ready for g4mi_set_wake.  wres, wgeo, wrou
 * (nslice*sample each, may be NULL) receive the single-particle wakes. */
int g4mi_wake_loss(int nslice, int sample, double ds, const double *current, const double *loss_ext,
                   double radius, int roundpipe, double conductivity, double relaxation, double gap,
                   double lgap, double hrough, double lrough, double *eloss, double *wres, double *wgeo,
                   double *wrou);

/* Field::initSolver (src/Core/Field.cpp:72-83) with
 * FieldSolverFFT::initSourceFilter (src/Core/FieldSolverFFT.cpp:159-170):
 * fft = 0 the ADI solver (default), 1 the spectral solver; filter = 1 with
 * xcut, ycut, sigmoid > 0 filters the source by 1/(1 + exp((r - 1)/sigmoid)),
 * r the radius in units of (xcut, ycut) of the normalised frequency.  The
 * spectral solver runs in the FP64 form only (G4MI_E_ARG otherwise). */
int g4mi_set_field_solver(g4mi_ctx *ctx, int fft, int filter, double xcut, double ycut, double sigmoid);

/* Control::applySlippage (src/Core/Control.cpp:158-276) on one device:
 * ring-offset rotation and zeroing of the inflowing slice. */
int g4mi_slippage(g4mi_ctx *ctx, double slip);

/* One step of Gencore::run (src/Core/Gencore.cpp:118-147): beam track,
 * field track, slippage. */
int g4mi_step(g4mi_ctx *ctx, const g4mi_step_params *st, double gammaref);

/* Several consecutive steps with one diagnostics record after each
 * (records array of nsteps entries, may be NULL); stops at a step whose
 * marker has bit 8 set, as Undulator::advance (src/Core/Undulator.cpp:112-139).
 * Returns the number of steps done (>= 0) or an error code. */
int g4mi_run(g4mi_ctx *ctx, const g4mi_step_params *steps, int nsteps, double gammaref,
             g4mi_diag_out *records);

/* Diagnostics of the current state (synchronises the stream). */
int g4mi_diag(g4mi_ctx *ctx, g4mi_diag_out *out);

/* Sorting::localSort target index (src/Util/Sorting.cpp:96) for every
 * particle: atar = floor(theta * (1/slen)); returns the count of particles
 * with atar != 0 in *moved. atar is a host array of total-particle length. */
int g4mi_sort_index(g4mi_ctx *ctx, double slen, int *atar, long long *moved);

/* No counterpart in the reference: g4mi_sort_index for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops).
 * Same inputs, outputs and results; it calls no allocator: the targets are
 * staged and counted in the library's static device scratch (16 MiB per
 * GPU, held until the process ends), so the beam may hold at most 4194304
 * particles (larger beams get G4MI_E_ARG).  Calls that overlap on one GPU
 * share that scratch and race. */
int g4mi_sort_index_noalloc(g4mi_ctx *ctx, double slen, int *atar, long long *moved);

/* Kernel schedule.  The first field step of each step size times every
 * ADI kernel eligible for its grid and step (two-kernel Thomas, fused
 * recurrence with and without prefetch, banded inverse in 32x32 tiles and
 * as a whole-width stream) on a scratch copy of its field and keeps the
 * fastest for that step size (the library's autotuner).  g4mi_set_schedule
 * fixes an entry instead ("adi_kernel": 0..4, -1 re-enables tuning; a
 * kernel not eligible for a step size falls back to the tuner's choice);
 * g4mi_get_schedule writes the current choice
 * and the measured timings as a JSON object into buf.  "fp64_fused" (1 on,
 * 0 off, -1 the default: on for windows of 256 slices or more): the step's
 * particle passes as two fused kernels (one workgroup per slice) where the
 * step allows it; smaller windows spread each slice over blocks of 256
 * particles in the separate kernels.  The spectral solver's first step likewise times its
 * complex-GEMM variants ("fft_variant": 0 four real products per complex
 * one and k chunks of 16, 1 three products, 2 four products with chunks of
 * 32, 3 three with 32; -1 tunes again) and its propagation by those GEMMs
 * against FFT convolutions of every row and column (zero-padded
 * power-of-two FFTs, ngrid <= 256), and keeps the faster
 * ("fft_propagation": 0 GEMMs, 1 FFT convolution, -1 tunes again).  The
 * source filter (with the FFT convolution) and the spectral diagnostics
 * (unless "fft_propagation" is 0) compute their DFTs as Bluestein chirp
 * convolutions on the same FFTs for ngrid <= 256.
 * Mixed form: its first 18 steps time the main kernel's shapes (4, 6 or 8
 * wavefronts per slice x 68 or 44 KiB of LDS, 4 or 6 x 32 KiB), the three ADI kernels,
 * and the step as a graph of kernels against one persistent
 * cooperative launch, and keep the fastest.  Entries: "adi_refine" (FP64
 * residual correction, default 1), "mixed_adi" (2 whole-width streaming,
 * 1 streaming strips, 0 tiles),
 * "main_nt" (256, 384, 512), "main_lds" (0: 68 KiB, 1: 44 KiB, 2: 32 KiB;
 * 512 threads take 44 KiB for 2),
 * "persistent", "graphs" (0/1); setting any of the kernel entries turns the
 * tuner off, "autotune" 1 turns it on again.  "mixed_fuse_front" (default
 * 1): within one g4mi_run the main kernel of each step but the last also
 * runs the next step's front (first transverse half step, slice sizes,
 * centroid, largest radius, footprint) for its slice, so the front kernel
 * of that step only takes over the slice sizes; 0 runs every front as its
 * own kernel. */
int g4mi_set_schedule(g4mi_ctx *ctx, const char *name, int value);
int g4mi_get_schedule(g4mi_ctx *ctx, char *buf, int len);

/* Wait for all work queued on the context. */
int g4mi_sync(g4mi_ctx *ctx);

/* Per-kernel timing with HIP events recorded on the context's stream into
 * a pool (no host synchronisation inside a step).  g4mi_set_timing(1)
 * starts recording and clears the totals; g4mi_kernel_times synchronises,
 * folds the recorded events and returns the average device milliseconds per
 * step of each phase: ms[0] transverse half steps (both), ms[1] space charge
 * (moments, long range, short range), ms[2] longitudinal push, ms[3] source
 * deposition, ms[4] ADI field advance (x and y sweeps), ms[5] slippage and
 * diagnostics.  nslots <= G4MI_NTIMERS. */
#define G4MI_NTIMERS 6
int g4mi_set_timing(g4mi_ctx *ctx, int on);
int g4mi_kernel_times(g4mi_ctx *ctx, double *ms, int nslots, long long *nsteps);

/* Text of the last error on this thread. */
const char *g4mi_last_error(void);

/* ---- input generation (host only, no GPU needed) ---------------------- */

/* Beam parameters of one slice, the BeamSlice record of the quiet loader
 * (include/QuietLoading.h:16-35 in the reference). */
typedef struct {
    double gamma, delgam, current, ex, ey, xcen, ycen, pxcen, pycen;
    double betax, betay, alphax, alphay, bunch, bunchphase, emod, emodphase;
} g4mi_slice_beam;

/* Parse a lattice file's text (LatticeParser::parse,
 * src/Lattice/LatticeParser.cpp:51-240) and unroll beamline `line` into
 * integration steps (Lattice::generateLattice, src/Lattice/Lattice.cpp:53-197).
 * Writes at most `max_steps` records; *nsteps receives the full count. */
int g4mi_parse_lattice(const char *text, const char *line, double delz, double lambda0,
                       double gamma0, g4mi_step_params *out, int max_steps, int *nsteps);

/* Generate the macroparticles of slices [0, nslice) as LoadBeam::init
 * (src/Loading/LoadBeam.cpp:127-212) does for the non-one4one case: quiet
 * Hammersley loading of each slice (global slice index islice0 + j), then,
 * for time-dependent runs with shot noise, one sequential ran2 stream of
 * ShotNoise (src/Loading/ShotNoise.cpp).  Output SoA arrays of off[nslice]
 * particles; per-slice counts must be multiples of nbins. */
int g4mi_load_beam(int nslice, const long long *off, const g4mi_slice_beam *slices, int nbins,
                   int shotnoise, int seed, int islice0, double lambda0, double sample,
                   int timerun, double *gamma, double *theta, double *x, double *y, double *px,
                   double *py);

/* LoadBeam::init (src/Loading/LoadBeam.cpp:127-212) straight into the
 * context's device arrays: the same beam as g4mi_load_beam followed by
 * g4mi_beam_upload with current[j] = slices[j].current, for the slice
 * counts the context was created with.  Quiet loading
 * (QuietLoading::loadQuiet, src/Loading/QuietLoading.cpp:50-260) runs on
 * the device, one workgroup per slice; the shot-noise draws
 * (ShotNoise::applyShotNoise, src/Loading/ShotNoise.cpp:29-68) come from the
 * host's sequential ran2 stream and are applied on the device.  Agrees with
 * the host loader to rounding (device log/exp/sin and block sums). */
int g4mi_load_beam_device(g4mi_ctx *ctx, const g4mi_slice_beam *slices, int nbins, int shotnoise,
                          int seed, int islice0, double lambda0, double sample, int timerun);

/* one4one beams (one macroparticle per electron), LoadBeam::init with
 * one4one (src/Loading/LoadBeam.cpp:129-132, 157-199): counts[j] =
 * round(current lambda0 sample / ce) electrons per slice; then every
 * coordinate of every slice, in slice order, from one ran2 stream seeded
 * with round(1e9 u), u the first draw of ran2(seed) (QuietLoading::init,
 * src/Loading/QuietLoading.cpp:15-40), no beamlets, theta scaled by 2 pi
 * sample, no shot noise. */
int g4mi_one4one_counts(int nslice, const g4mi_slice_beam *slices, double lambda0, double sample,
                        long long *counts);
int g4mi_load_beam_one4one(int nslice, const long long *off, const g4mi_slice_beam *slices, int seed,
                           double sample, double *gamma, double *theta, double *x, double *y, double *px,
                           double *py);

/* Gaussian seed field of one slice (GaussHermite::loadGauss,
 * src/Util/GaussHermite.cpp:20-63, fundamental TEM00), interleaved complex. */
int g4mi_load_gauss(double *field, int ngrid, double gridmax, double lambda, double power,
                    double waist, double waist_pos, double phase);

/* Whole input deck: parse the namelist file (src/Main/GenMain.cpp:80-494
 * subset: &setup &time &field &beam &efield &track) and lattice, generate
 * beam and field, run every &track on the GPU and write <rootname>.out.h5
 * in the reference layout (src/IO/Output.cpp) when HDF5 is available.
 * Replaces genmain() for the supported commands. */
int g4mi_run_deck(const char *input_file, int device, int quiet);
/* The same deck without a device: parses it, matches the lattice, generates
 * beam and field and, when &setup names a rootname, writes the output file
 * with the /Meta, /Global, /Lattice and once-per-run /Beam and /Field
 * datasets.  `json` receives a summary of the last &track (slices,
 * macroparticles, grid, steps, output records, matched Twiss).  No
 * counterpart in the reference: the setup half of genmain() alone. */
int g4mi_check_deck(const char *input_file, char *json, int len);
/* A double dataset of an HDF5 file (rank 1 or 2): up to `cap` values into
 * `out`, its shape into dims[2] (dims[1] = 0 for rank 1); returns the
 * element count.  Reads the output files back without h5py. */
long long g4mi_h5_read(const char *file, const char *dataset, double *out, long long cap,
                       long long *dims);

/* ------------------------------------------------------------------------
 * Learned field step (no counterpart in the reference).  A Fourier neural
 * operator maps a slice's field E and deposited source S to the field
 * after the ADI step; it is trained on (E, S, E_next) triplets captured
 * from the solver's own steps and can stand in for the ADI of a mixed-form
 * context, with the ADI as fallback.  Every dense product runs on BF16
 * MFMA with FP32 accumulation; weights and Adam state are FP32.
 * ------------------------------------------------------------------------ */
typedef struct g4mi_fno g4mi_fno;

/* width channels, modes x modes retained Fourier modes (even), layers
 * Fourier layers, hidden projection channels, max_batch slices per pass.
 * width and hidden are at most 256. */
int g4mi_fno_create(g4mi_fno **fno, int device, int ngrid, int width, int modes, int layers,
                    int hidden, int max_batch, unsigned seed);
int g4mi_fno_destroy(g4mi_fno *fno);
long long g4mi_fno_param_count(g4mi_fno *fno);
/* Parameter vector, in order: lift P[4][width], bP[width]; per layer
 * Rr[modes^2][width][width], Ri[...], W[width][width], b[width];
 * Q1[width][hidden], bQ1[hidden], Q2[hidden][2], bQ2[2]. */
int g4mi_fno_get_params(g4mi_fno *fno, float *params);
int g4mi_fno_set_params(g4mi_fno *fno, const float *params);
/* in [batch][ngrid^2][4] = (Re E, Im E, Re S, Im S) / scale,
 * out [batch][ngrid^2][2] = (E_next - E) / scale (the network output). */
int g4mi_fno_forward(g4mi_fno *fno, const float *in, float *out, int batch);
/* Mean-squared loss against target [batch][ngrid^2][2] and its gradient
 * with respect to every parameter (same order as g4mi_fno_get_params). */
int g4mi_fno_loss_grad(g4mi_fno *fno, const float *in, const float *target, int batch,
                       double *loss, float *grad);
/* Training set on the device: capacity, pairs held. */
int g4mi_fno_reserve(g4mi_fno *fno, long long npairs);
long long g4mi_fno_pair_count(g4mi_fno *fno);
/* Adam with a cosine learning-rate schedule over `epochs`; the last
 * `nholdout` pairs are kept out of training; per-epoch mean losses. */
int g4mi_fno_train(g4mi_fno *fno, int epochs, int batch, float lr, long long nholdout,
                   unsigned seed, double *train_loss, double *holdout_loss);
/* Epochs [first_epoch, first_epoch + epochs) of a schedule of total_epochs
 * (the same cosine schedule, split over several calls); losses per epoch
 * of this call. */
int g4mi_fno_train_span(g4mi_fno *fno, int first_epoch, int epochs, int total_epochs, int batch,
                        float lr, long long nholdout, unsigned seed, double *train_loss,
                        double *holdout_loss);
int g4mi_fno_eval(g4mi_fno *fno, long long first, long long last, int batch, double *loss);

/* Attach a network to a mixed-form (G4MI_FP32) context:
 *   mode 0: detach;
 *   mode 1: capture -- after every step's ADI, the (E, S, E_next) pairs of
 *           every `every`-th beam slice are appended to the training set;
 *   mode 2: replace -- the network advances the field; the ADI runs
 *           instead on every `every`-th step and whenever the
 *           particle-to-field energy-balance residual of the network's
 *           step exceeds `threshold`:
 *             r = |dP_field + sum_s I_s eev d<gamma>_s - 4 sum |S|^2 pw| / P_field,
 *           dP_field the change of the total field power over the step,
 *           d<gamma>_s the push's mean energy change of slice s and the
 *           last term the source's own energy (S the deposited source,
 *           pw the power per |E|^2 of a cell).  On the every-th steps the
 *           network's step is also computed and compared with the ADI's
 *           (g4mi_fno_judge). */
int g4mi_attach_fno(g4mi_ctx *ctx, g4mi_fno *fno, int mode, int every, double threshold);
/* Steps advanced by the network and by the ADI since attaching, and the
 * mean residual of the network's steps and of the ADI's steps. */
int g4mi_fno_stats(g4mi_ctx *ctx, long long *fno_steps, long long *adi_steps,
                   double *mean_residual_fno, double *mean_residual_adi);
/* Judge steps since attaching and the mean over them of the network's
 * relative field error ||E_fno - E_adi|| / ||E_adi|| and relative update
 * error ||E_fno - E_adi|| / ||E_adi - E|| (norms over all slices); and the
 * mean energy-balance residual of every network step, accepted or not. */
int g4mi_fno_judge(g4mi_ctx *ctx, long long *njudge, double *rel_field_err,
                   double *rel_update_err, double *mean_residual_all);

#ifdef __cplusplus
}
#endif
#endif
