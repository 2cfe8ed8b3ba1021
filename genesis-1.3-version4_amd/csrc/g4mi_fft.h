// Spectral field solver on the device (FieldSolverFFT, src/Core/FieldSolverFFT.cpp).
//
// A step multiplies the 2-D DFT of every field slice by exp(K2 delz) and
// transforms back.  The phase factor is separable in kx and ky, so the step
// is E' = M E M^T with M the circulant matrix of the 1-D propagator
// (M[j][l] = m[(j - l) mod n], m = inverse DFT of the 1-D phase / n): two
// complex products with one n x n matrix per slice, run as FP64 MFMA GEMMs.
// The source filter (FieldSolverFFT.cpp:72-87) is not separable and runs as
// four products with the DFT matrices and a pointwise sigmoid.
#pragma once

#include "g4mi_common.h"

namespace g4mi {

struct CgemmArgs {
    const double2 *a;    // row mode: data rows [nrows][ng]
    const double2 *b;    // column mode: data slices [ns][ng][ng]
    const double2 *mat;  // [ng][ng] row-major complex matrix
    double2 *c;          // output, the layout of the data operand
    long long nrows;     // row mode: rows of a
    int ng, ns;
    double scale;        // applied to the product
    // column mode epilogue: c = scale * ptw * (mat . b[s]) + 2 add[ii] + 2 src[ii] (in box[ii])
    const double *ptw;   // [ng][ng] real factor or null
    const double2 *add;  // [ns][ng][ng] or null
    const double2 *src;  // [ns][ng][ng], valid inside box[ii] only, or null
    const int4 *box;     // {ix0, iy0, ix1, iy1} per beam slice
    int mask_a;          // row mode: rows are source planes, zero outside box[row / ng]
    int first, ring;     // column mode: ii = ring ? (s - first + ns) % ns : s
    int variant;         // 0: four real products per complex one, k chunks of 16;
                         // 1: three products, 16; 2: four, 32; 3: three, 32
};
constexpr int kCgemmVariants = 4;

// row mode: c[r][j] = scale * sum_l a[r][l] mat[j][l]          (a . mat^T)
// column mode: c[s][j][x] = scale * sum_l mat[j][l] b[s][l][x] (+ epilogue)
int launch_cgemm_rows(hipStream_t s, const CgemmArgs &g);
int launch_cgemm_cols(hipStream_t s, const CgemmArgs &g);
// output tile edge used for ngrid (32, 48, 64 or 80)
int cgemm_tile(int ng);

// The propagation E' = M E M^T (+ 2 S or 2 Y) as circular convolutions of
// every line with the propagator kernel, each by a zero-padded FFT of length
// L (Stockham radix-R passes in LDS, one wavefront per line, 8 lines per
// workgroup): rows pass out = in M^T over all ns*ng rows; columns pass
// out = M in per slice with the column epilogue of the GEMM (add, src in
// box, ring).  hk/tw: the tables of fft_conv_tables, L entries each.
// With pre/post chirps (fft_chirp_tables) a pass is a length-n DFT of every
// line instead (Bluestein): each line is multiplied by pre before and by
// post after the convolution.
struct ConvArgs {
    const double2 *in;
    double2 *out;
    const double2 *hk, *tw;
    long long nrows;  // rows pass: ns * ng
    int ng, ns, L;
    const double2 *pre, *post;  // [ng] per-position factors, or null
    int mask;                   // rows pass: rows are source planes, zero outside box[row / ng]
    // columns pass epilogue, as CgemmArgs: out = scale * ptw * result + 2 add + 2 src (in box)
    double scale = 1.0;
    const double *ptw;
    const double2 *add;
    const double2 *src;
    const int4 *box;
    int first, ring;
};
int launch_fft_conv_rows(hipStream_t s, const ConvArgs &a);
int launch_fft_conv_cols(hipStream_t s, const ConvArgs &a);

// Moments of |F|^2 over the FFT index of each of ns slices F[ns][ng][ng]
// (DiagField::getValues, src/Core/Diagnostic.cpp:789-807): out[5 s + 0..4] =
// sum w, sum dx w, sum dx^2 w, sum dy w, sum dy^2 w, with w = |F|^2 and
// (dx, dy) the centred offsets of the grid point whose FFT index it is.
// part: fft_moments_parts(ns) * 5 ns doubles of scratch (at most
// kFftMomentParts blocks per slice, so that a window of few slices still
// fills the device).
constexpr int kFftMomentParts = 64;
int fft_moments_parts(int ns);
void launch_fft_moments(hipStream_t s, int ns, const double2 *F, int ng, double *part, double *out);
// FP32 field -> FP64 copy (n complex values)
void launch_widen_field(hipStream_t s, const float2 *in, double2 *out, long long n);

}  // namespace g4mi
