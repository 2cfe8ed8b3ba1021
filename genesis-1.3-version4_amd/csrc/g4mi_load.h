// Device beam loading (quiet Hammersley loading, shot-noise kicks and the
// conversion to the mixed form's FP32 offsets), kernels in g4mi_load.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "g4mi.h"

namespace g4mi {

// Quiet loading of every slice, one workgroup per slice: slice is gets
// off[is+1]-off[is] particles (a multiple of nbins) from the Hammersley
// sequences that start after starts[is]; theta scaled by theta0.
void launch_quiet_load(hipStream_t s, int ns, const long long *off, const unsigned int *starts,
                       const g4mi_slice_beam *slices, int nbins, double theta0, double *g, double *th, double *x,
                       double *y, double *px, double *py);
// th[k] += -sum_h an sin((h+1) th[k] + phi), (phi, an) pairs of slice is at
// draws + 2 (dbase[is] + h mpart + beamlet); dbase[is] < 0 skips the slice.
void launch_shot_noise(hipStream_t s, int ns, const long long *off, const long long *dbase, int nbins, int nh,
                       const double *draws, double *th, int maxpart);
// base[is] = 2 pi floor(<th>_slice / 2 pi)
void launch_theta_base(hipStream_t s, int ns, const long long *off, const double *th, double *base);
// out[b] = partial sums of g[0, n) over nblocks workgroups
void launch_sum_gamma(hipStream_t s, const double *g, long long n, double *out, int nblocks);
// FP32 offsets of the mixed form: g - gref, th - base[slice], the rest rounded
void launch_to_offsets(hipStream_t s, int ns, const long long *off, long long n, const double *base, double gref,
                       const double *g, const double *th, const double *x, const double *y, const double *px,
                       const double *py, float *fg, float *ft, float *fx, float *fy, float *fpx, float *fpy,
                       int maxpart);

}  // namespace g4mi
