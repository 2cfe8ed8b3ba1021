// Particle migration of one4one beams (Beam::sort) on the device.
#pragma once

#include "g4mi_common.h"

namespace g4mi {

// Work buffers of one sort of n particles: keys/values in and out, the
// corrected phases, the radix sort's digit counts and scan totals, and the six
// destination coordinate arrays (swapped with the beam's after the sort).
struct SortWork {
    int *keys, *keys_out, *vals, *vals_out;
    double *th_new;
    void *temp;
    size_t temp_bytes;
    double *g, *th, *x, *y, *px, *py;
};

// storage the radix sort needs for n keys in 0..ns: 16 digit counts per
// tile of 2048 keys and their scan's chunk totals (the storage for the
// initial particle count serves every later, smaller sort)
int sort_work_bytes(int ns, long long n, size_t *bytes);
// Sort the beam (off[ns+1] on the device, n particles) into w's coordinate
// arrays; the new offsets go to off_new[ns+1] (device).
int launch_sort_one4one(hipStream_t st, int ns, long long n, const long long *off, double slen, double *g,
                        double *th, double *x, double *y, double *px, double *py, const SortWork &w,
                        long long *off_new);

}  // namespace g4mi
