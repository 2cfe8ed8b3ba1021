// Launchers of the FP64 device kernels (defined in g4mi_kernels.hip).
#pragma once

#include "g4mi_common.h"

namespace g4mi {

// Thomas factors of the constant-coefficient ADI systems (complex, ngrid
// entries each, device pointers) and cstep = (0, cstep).
struct AdiFactors {
    const double2 *c, *cbet, *cwet;
    double cstep;
    int ngrid;
};

void launch_transverse(hipStream_t s, long long n, const double *g, double *x, double *y,
                       double *px, double *py, const TransParams &p);
void launch_r56(hipStream_t s, long long n, const double *g, double *th, double r56, double g0);
// Energy kicks after the push (Beam::track, src/Core/Beam.cpp:105-106), in
// the reference's order: gamma += beamlet_dg[boff[s] + (i - off[s]) / nbins]
// (Incoherent::apply) when beamlet_dg is set, then gamma += slice_dg[s]
// (Collective::apply) when slice_dg is set.  One workgroup per slice.
void launch_kicks(hipStream_t s, int ns, const long long *off, double *g, const double *beamlet_dg,
                  const long long *boff, int nbins, const double *slice_dg);
void launch_slice_moments(hipStream_t s, int ns, const long long *off, const double *x,
                          const double *y, double *mom);
void launch_slice_sizes(hipStream_t s, int ns, const long long *off, const double *cur,
                        const double *mom, double *fsize, double *cen);
void launch_lsc_long(hipStream_t s, int ns, const double *cur, const double *fsize, double sl,
                     double gamma, double scl, double *esc);
void launch_sc_rbound(hipStream_t s, int ns, const long long *off, const double *x,
                      const double *y, const double *cen, double *rbound);
void launch_sc_scan(hipStream_t s, int ns, const long long *off, const double *rbound,
                    double *rmax_state, int ngr, double *dr);
void launch_sc_solve(hipStream_t s, int ns, const long long *off, const double *x,
                     const double *y, const double *th, const double *cur, const double *cen,
                     const double *dr, int ngr, int nz, int nphi, double ks, double gz2,
                     double *ez, double *efield);
void launch_push(hipStream_t s, int nblocks, const int *blk_slice, const long long *blk_begin,
                 const long long *off, double *g, double *th, const double *x, const double *y,
                 const double *px, const double *py, const double2 *field, const double *esc,
                 const double *ezsc, const PushParams &p);
// The step's particle passes in two kernels (one workgroup per slice).
// front: first transverse half step, moments mom[4 ns] (the same bits as
// launch_slice_moments), fsize, cen, rbound (when sc), the in-grid
// footprint box pbox[ns] of each slice and pmaxf[ns], the largest
// sqrt(faw2) / gamma of its in-grid particles (the deposit's fixed-point
// scale).
void launch_fp64_front(hipStream_t s, int ns, const long long *off, const double *g, double *x, double *y,
                       double *px, double *py, const double *cur, const TransParams &t1, bool moments, bool sc,
                       double gridmax, double dgrid, int ng, double *mom, double *fsize, double *cen,
                       double *rbound, int4 *pbox, const DepositParams &d, double *pmaxf);
// Parameters of the main kernel: the push (fundamental only), the second
// transverse half step, R56, the short-range solver and the deposit.
struct FuseParams {
    PushParams p;
    TransParams t2;
    DepositParams d;
    double r56, gamma0;            // R56 (chicane != 0)
    double ks, gz2;                // short-range solver
    int chicane, use_sc, ngr, nz, nphi;
};
// Mode-solution cells the main kernel holds: (2 nphi + 1) nz ngr must fit.
int fp64_main_modes();
// main: short-range solve, push, R56, second transverse half step, beam sums
// beamsum[8 ns] (the layout of launch_diag_beam) and the deposit of the
// slice's source (src, box) when f.d.enabled.
void launch_fp64_main(hipStream_t s, int ns, const long long *off, double *g, double *th, double *x, double *y,
                      double *px, double *py, const double2 *field, const double *esc, const double *cur,
                      const double *cen, const double *dr, double *efield, const int4 *pbox,
                      const double *pmaxf, double2 *src, int4 *box, double *beamsum, const FuseParams &f);
void launch_deposit(hipStream_t s, int ns, const long long *off, const double *g,
                    const double *th, const double *x, const double *y, const double *cur,
                    double2 *src, int4 *box, const DepositParams &p);
void launch_adi(hipStream_t s, int ns, double2 *field, double2 *work, const double2 *src,
                const int4 *box, const AdiFactors &f, int first);
// Fused two-sweep ADI (one workgroup per slice, ngrid <= 256) with
// register-window back substitution of `tile` cells (8 or 16), tiles
// prefetched into registers (prefetch=1) or loaded when needed (0), and the
// field diagnostics sums of every storage slice into diag[8*ns].
int launch_adi_fused(hipStream_t s, int ns, double2 *field, double2 *work, const double2 *src,
                     const int4 *box, const AdiFactors &f, int first, double *diag, int tile,
                     int prefetch);
// Banded-inverse ADI (both half steps, 32x32 output tiles, halo M+1):
// reads field `fin`, writes `fout` (the host swaps the buffers), band =
// [ngrid][2M+1] complex rows of A^-1 around the diagonal, a = Im(cstep).
// m: 7, 8 or 12.  Rows [klo, khi] of the band equal row (klo+khi)/2 to
// 2^-60.  tilediag: scratch [ns][adi_band_tiles(ngrid)][8].
int adi_band_tiles(int ng);
int launch_adi_band(hipStream_t s, int ns, const double2 *fin, double2 *fout, const double2 *src,
                    const int4 *box, const double2 *band, double a, int ng, int first, int m,
                    int klo, int khi, double *tilediag, double *diag);
// The banded operator as a whole-width stream (one workgroup per slice, one
// lane per column; the products of launch_adi_band as four partial sums,
// diagnostics straight into diag[8 ns]).  Eligible for band half widths 7 and 8, grids up to 512
// columns whose rows [klo, khi] share the central band except in the first
// and last 32 columns; returns -1 otherwise.
bool adi_wide64_ok(int ng, int m, int klo, int khi);
int launch_adi_wide64(hipStream_t s, int ns, const double2 *fin, double2 *fout, const double2 *src,
                      const int4 *box, const double2 *band, double a, int ng, int first, int m, int klo, int khi,
                      double *diag);
void launch_zero(hipStream_t s, double2 *p, long long n);
// n doubles from device memory to pinned host memory (dst: its device address)
void launch_copy_out(hipStream_t s, const double *src, double *dst, long long n);
// harmonic bunching sums for h0 .. h0 + cnt - 1 (cnt <= 8), out[ns][cnt][2]
void launch_diag_bunch(hipStream_t s, int ns, const long long *off, const double *th, int h0, int cnt,
                       double *out);
void launch_diag_beam(hipStream_t s, int ns, const long long *off, const double *g,
                      const double *th, const double *x, const double *y, double *out);
// px/py sums, coordinate extremes and the on-axis field value per beam
// slice, out[16 ns]
void launch_diag_aux(hipStream_t s, int ns, const long long *off, const double *g, const double *x,
                     const double *y, const double *px, const double *py, const double2 *field, int ng,
                     int first, double *out);
void launch_diag_field(hipStream_t s, int ns, const double2 *field, int ng, double *out);
void launch_sort_index(hipStream_t s, long long n, const double *th, double invslen, int *atar,
                       unsigned long long *moved);

}  // namespace g4mi
