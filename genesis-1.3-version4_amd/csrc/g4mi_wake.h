// Wakefield set-up on the host (Wake, src/Main/Wake.cpp; Collective,
// src/Core/Collective.cpp).  All arrays are FP64; `ns` counts slices at the
// highest resolution (simulation slices times the sample rate), `ds` is
// their length.
#pragma once

#include <vector>

namespace g4mi {

// single-particle wakes at s = i ds, i < ns
void wake_resistive(int ns, double ds, double radius, double conductivity, double relaxation, bool roundpipe,
                    double *w);
void wake_geometric(int ns, double ds, double radius, double gap, double lgap, bool roundpipe, double *w);
void wake_roughness(int ns, double ds, double radius, double hrough, double lrough, double *w);
// Loss per simulation slice [eV/m]: ext (the `loss` input) plus the
// convolution of the current profile cur[ncur] (A) with the wakes
// (ns = ncur * sample_hi entries each).
void wake_loss(int ncur, int sample_hi, double ds, const double *cur, const double *ext, const double *wres,
               const double *wgeo, const double *wrou, double *eloss);

}  // namespace g4mi
