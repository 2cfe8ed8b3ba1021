// Mixed-precision form of the undulator step (FP32 particles and field).
//
// Particle state in HBM, per macroparticle, 24 bytes:
//   dg  = gamma - gref            (FP32; gref is the context's FP64 reference energy)
//   dth = theta - base[slice]     (FP32; base is an FP64 per-slice phase)
//   x, y, px, py                  (FP32)
// The FP64 slice base advances every step by kappa*dz, the phase advance of
// a reference particle (gamma = gref, on axis, no field), so the FP32
// offsets only integrate each particle's departure from it.  The field is
// complex FP32 in two buffers (the ADI writes the other one; `cur` in the
// device state names the live buffer).
//
// Every per-step parameter sits in a device-resident step table and every
// kernel reads steps[state->istep], so one captured graph per band width
// replays any step of a run with no host work in between.
#pragma once

#include "g4mi_common.h"

namespace g4mi {

constexpr int kMixMaxZero = 64;     // slices slippage may zero in one step
constexpr int kMixRecord = 16;      // doubles per slice and step in the record buffer

// One integration step of the mixed form, prepared on the host.
struct MixStep {
    TransParams t1, t2;             // first / second transverse half step
    PushParams p;                   // p.first / p.nslice unused (device state)
    DepositParams d;
    double gamma0;                  // reference energy of the step (R56, LSC)
    double r56;                     // chicane R56 (chicane != 0)
    double kappa;                   // d(base)/dz
    double eps_ref;                 // (1 + aw^2) / gref^2 of the reference particle
    double lsc_ds, lsc_scl;         // long-range space charge (EFieldSolver::longRange):
                                    // ds per slice = slice length * gamma, scale
    double sc_gz2, sc_ks;           // short-range space charge (EFieldSolver::shortRange)
    double slip;                    // slippage of the step in slices
    int longrange, use_sc, chicane, band;
    int fnext;                      // the main kernel also runs the next step's front for its slice
    int front_done;                 // the previous step's main kernel ran this step's front
    int fbuf;                       // the slice-size buffer of this step (mom_fsize[fbuf])
};

// Band of the FP32 inverse for one step size.
struct MixBand {
    long long off;                  // offset of [ng][W] float2 rows in the band table
    double a;                       // rtmp = dz / (4 ks dgrid^2)
    int klo, khi;                   // rows sharing the band of (klo+khi)/2
    int m;                          // half width the step size needs (<= template M)
    int pad;
};

// Device-resident run state.
struct MixState {
    int istep;                      // step-table index of the step being run
    int rec_step;                   // step index of the last completed step
    int first;                      // ring offset of the field (Control.cpp:205-274)
    int cur;                        // live field buffer (0/1)
    double accuslip;
    int nzero;
    int zero[kMixMaxZero];
    unsigned seq;                   // steps completed by this context (rounding draws)
};

struct MixArrays {
    // particles
    float *dg, *dth, *x, *y, *px, *py;
    float *xs, *ys;                 // post-t2 positions when the next front rides in the main loop
    double *base;                   // [ns]
    const long long *off;           // [ns+1]
    const double *cur;              // [ns] current
    // field
    float2 *field[2];
    float2 *src;                    // [ns][ng*ng] source (inside box only)
    int4 *box;                      // [ns] box of the source plane the ADI reads
    int4 *pbox;                     // [ns] cell footprint after the first half step
    // per-slice scratch
    // slice sizes, two buffers: step S reads mom_fsize[S.fbuf] while its main
    // kernels write the next step's into the other one
    double *mom_fsize[2];
    double *cen, *rbound, *dr, *rmax, *esc, *efield;
    double *beamsum;                // [ns][8]
    double *egsum;                  // [ns] sum of the push's gamma increments (before rounding)
    double *tilediag;               // [ns][tiles][8]
    double *rec;                    // [nrec][ns][kMixRecord]
    int nrec;
    // tables
    const MixStep *steps;
    MixState *st;
    const float2 *bands;
    const MixBand *binfo;
    // constants
    int ns, ng, sc_ng, sc_nz, sc_nphi;
    double gref, gridmax, dgrid, sample;
    int timerun, periodic;
};

// Kernel attributes (dynamic LDS above 64 KiB), set once per process before
// any launch or graph capture; 0 on success.
int mix_prepare();
// LDS cells the main kernel has for the space-charge mode solutions
// ((2 nphi + 1) nz modes x radial cells).
int mix_sc_mode_cells();

// Band half widths the mixed ADI is instantiated for.
bool mix_band_supported(int m, int *tmpl_m);
int mix_band_tiles(int ng);
size_t mix_band_lds(int tmpl_m, bool refine);

// One step: front (transverse + moments + radius), radial scan, main
// (space charge + push + R56 + transverse + deposit + beam sums), ADI
// (template band width tmpl_m, FP64 residual correction when refine),
// slippage, zeroing and the step record.  ev (optional, 9 events) brackets
// the kernel groups: front [0,1], main [3 - 2], ADI [6,7], slippage + record
// [7,8]; the other intervals are empty.
// adi_mode: kAdiTile (32x32 tiles, correction up to a half width of 9),
// kAdiStrips (streaming, one wavefront per column strip) or kAdiWide
// (streaming, one workgroup per slice covering all columns; grids up to 512
// columns, wider ones run the strips).
// main_nt: threads per slice of the main kernel (256, 384, 512); main_lds: its
// dynamic LDS region, 0 = 68 KiB, 1 = 44 KiB.
constexpr int kAdiTile = 0, kAdiStrips = 1, kAdiWide = 2;
int launch_mix_step(hipStream_t s, const MixArrays &a, int tmpl_m, bool refine, bool sc,
                    bool moments, int maxzero, hipEvent_t *ev, int adi_mode, int main_nt, int main_lds,
                    bool front = true);
// The same step as one cooperative launch of co-resident workgroups with
// grid barriers between the phases (the streaming ADI); ev (optional, 9
// events): [0] before, [8] after.  Returns -1 if the launch is refused.
int launch_mix_step_persistent(hipStream_t s, const MixArrays &a, int tmpl_m, bool refine, bool sc, bool moments,
                               int maxzero, hipEvent_t *ev);
// The same step in three parts: the beam kernels (front, scan, main; ev
// slots 0-6), the ADI (ntiles: the field-sum blocks per slice it wrote) and
// the tail (slippage, zeroing, record of ntiles blocks; ev slots 7-8).
// egsum: also sum the push's energy increments per slice (a.egsum)
// front: launch the front kernel (false when the previous step's main kernel ran it)
void launch_mix_beam(hipStream_t s, const MixArrays &a, bool sc, bool moments, hipEvent_t *ev, bool egsum,
                     int main_nt = 384, int main_lds = 0, bool front = true);
int launch_mix_adi(hipStream_t s, const MixArrays &a, int tmpl_m, bool refine, int adi_mode, int *ntiles);
void launch_mix_tail(hipStream_t s, const MixArrays &a, int maxzero, int ntiles, hipEvent_t *ev);
// Field sums per storage slice ([ns][8], as the ADI's diagnostics with one
// block per slice) of the live buffer (next = 0) or the other one (1).
void launch_mix_field_sums(hipStream_t s, const MixArrays &a, int next, double *out8);
// sum of |S|^2 over each beam slice's source box, out[ns]
void launch_mix_src_sums(hipStream_t s, const MixArrays &a, double *out);

// Standalone sums for a diagnostics call outside a run:
// beam [ns][8] and field per storage slice [ns][8].
void launch_mix_diag_aux(hipStream_t s, const MixArrays &a, double *out);
void launch_mix_diag(hipStream_t s, const MixArrays &a, double *beam8, double *field8);
void launch_mix_sort_index(hipStream_t s, const MixArrays &a, double invslen, int *atar,
                           unsigned long long *moved);

}  // namespace g4mi
