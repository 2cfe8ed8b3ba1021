// Host-side numerics of the g4mi library (compiled with g++, no HIP).
//
// These are the per-step scalars the device kernels consume: undulator
// coupling, ADI Thomas factors, chicane matrices, and the input generators
// (lattice, beam, field).  They run once per step or once per run on the
// host, in FP64 with the reference's operation order.
#pragma once

#include <string>
#include <vector>

#include "g4mi.h"

namespace g4mi {

// Charge of the electron in the reference's units (src/Main/GenMain.cpp:68).
inline double kCeH() { return 4.8032045e-11; }

// Undulator::fc (src/Core/Undulator.cpp:141-167) for harmonic h.
double coupling(const g4mi_step_params &st, int h);
// BesselJ::value (src/Util/BesselJ.cpp:101-159).
double bessel_j(int n, double x);

// FieldSolverADI::init (src/Core/FieldSolverADI.cpp:125-182): interleaved
// complex arrays c, cbet, cwet of ngrid entries; returns Im(cstep).
double adi_factors(int ngrid, double delz, double dgrid, double xks, double *c, double *cbet,
                   double *cwet);

// Band of the inverse of the ADI tridiagonal matrix: band[k][m] =
// (A^-1)[k][k+m-M] (interleaved complex, ngrid*(2M+1) entries, zero where
// k+m-M is off the grid), computed column by column with the Thomas factors
// of adi_factors.  Returns the largest entry outside the band relative to
// its row's diagonal entry.
double adi_band(int ngrid, const double *c, const double *cbet, const double *cwet, int M,
                double *band);

// Spectral solver matrices (FieldSolverFFT, src/Core/FieldSolverFFT.cpp),
// interleaved complex, row-major n x n.
// The step's propagator: M[j][l] = m[(j - l) mod n] with
// m[d] = (1/n) sum_k p_k exp(2 pi i k d / n) and p_k = exp(i K2_1(k) delz),
// K2_1(k) = -dx(k)^2 dk^2 / (2 xks) the one-dimensional part of K2 (:139)
// at FFT index k.  exp(K2 delz) = p_ky p_kx, so the 2-D step
// IDFT(exp(K2 delz) DFT(E)) / n^2 equals M E M^T.
void fft_propagator(int n, double dgrid, double xks, double delz, double *m);
// The same step as a circular convolution of every line with m, run as a
// zero-padded FFT of length L = fft_conv_length(n) (the power of two >= 2n-1,
// at least 128; 0 when n > 256): hk[k] = (1/L) sum_{|d|<n} m[d mod n]
// exp(-2 pi i k d / L) and tw[k] = exp(-2 pi i k / L), L interleaved complex
// entries each.  Returns L (0: not supported for this n).
int fft_conv_length(int n);
int fft_conv_tables(int n, double dgrid, double xks, double delz, double *hk, double *tw);
// A length-n DFT with exp(sign 2 pi i j k / n) as the same convolution
// (Bluestein): chirp[j] = exp(sign i pi j^2 / n) multiplies the line before
// and after, hk = (1/L) DFT_L of conj(chirp) placed at d mod L for |d| < n.
// n entries of chirp, L of hk (interleaved complex); returns L.
int fft_chirp_tables(int n, int sign, double *hk, double *chirp);
// W[j][l] = exp(sign 2 pi i j l / n), the exponent reduced mod n exactly.
void dft_matrix(int n, int sign, double *w);
// The source filter's sigmoid at FFT index [ky][kx] (:141-147), n*n reals.
void fft_sigmoid(int n, double xc, double yc, double sig, double *s);

// Chicane transfer matrix of TrackBeam::applyChicane (:131-221), row-major 4x4.
void chicane_matrix(double angle, double lb, double ld, double lt, double m[16]);
// TrackBeam::applyR56 factor (:260-269), already scaled to theta units.
double chicane_r56(double angle, double lb, double ld, double lambda0, double gamma0);

// ---- input generation (Loading/, Lattice/) ----
struct SliceBeam {  // QuietLoading.h BeamSlice
    double gamma, delgam, current, ex, ey, xcen, ycen, pxcen, pycen;
    double betax, betay, alphax, alphay, bunch, bunchphase, emod, emodphase;
};
// Sequential ran2 generator (src/Util/RandomU.cpp).
class Ran2 {
  public:
    explicit Ran2(unsigned int seed);
    double next();
    // the generator's whole state (shuffle table, last output, both LCGs)
    static constexpr int kStateInts = 35;
    void save(int *state) const;
    void load(const int *state);

  private:
    int iv_[32], iy_, s1_, s2_;
};
// Quiet loading of one slice (src/Loading/QuietLoading.cpp:50-260), SoA.
void quiet_load(double *gamma, double *theta, double *x, double *y, double *px, double *py,
                const SliceBeam &sl, int npart, int nbins, double theta0, int islice);
// one4one loading (QuietLoading::init/loadQuiet with one4one,
// src/Loading/QuietLoading.cpp:15-40, 50-260): every coordinate of every
// particle from one ran2 stream seeded with one4one_seed(seed, rank),
// which runs on from slice to slice; no beamlets.
unsigned int one4one_seed(int seed, int rank);
void quiet_load_one4one(Ran2 &stream, double *gamma, double *theta, double *x, double *y, double *px,
                        double *py, const SliceBeam &sl, int npart, double theta0);
// Shot noise (src/Loading/ShotNoise.cpp:29-68) consuming `stream`.
void shot_noise(Ran2 &stream, double *theta, int npart, int nbins, double ne);
// The draws of shot_noise for one slice in stream order, appended to out as
// (phase, amplitude) pairs, harmonic-major then beamlet; returns the pair
// count (0 when ne <= 0: the stream is then not advanced).
long long shot_noise_draws(Ran2 &stream, int npart, int nbins, double ne, std::vector<double> &out);
// First Hammersley index of slice islice's sequences: round(1e9 u), u the
// first ran2 draw seeded with the slice index (QuietLoading.cpp:54-65).
unsigned int quiet_start(int islice);
// Seed stream of ShotNoise::init (:16-27).
unsigned int shot_noise_seed(int seed, int rank);
double inverse_erfc(double y);
// Gaussian seed field (src/Util/GaussHermite.cpp:20-63), TEM00 at harmonic harm
// (wavenumber harm 2 pi / lambda).
void gauss_field(double *field, int ngrid, double gridmax, double lambda, double power,
                 double w0, double z0, double phase, double xcen, double ycen, double xangle = 0.0,
                 double yangle = 0.0, int harm = 1);

// Lattice file (LatticeParser.cpp) -> unrolled steps (Lattice.cpp).
struct LatticeResult {
    std::vector<g4mi_step_params> steps;
    std::string error;
};
LatticeResult parse_and_unroll(const std::string &text, const std::string &beamline, double delz,
                               double lambda0, double gamma0);

}  // namespace g4mi
