// The growing radial grid of the short-range space-charge solver
// (EFieldSolver.cpp:159-165): rmax is solver state that grows to 1.5 rbound
// whenever a slice exceeds it, visited in slice order.  One wavefront walks
// 64 slices at a time and a ballot finds the next slice that raises rmax, so
// the serial part is one step per growth event, not per slice.  The whole
// workgroup first stages the slices' radii in LDS, kScanChunk slices at a
// time, so the walk itself waits on no global loads.
#pragma once

#include <hip/hip_runtime.h>

namespace g4mi {

constexpr int kScanNT = 1024;      // threads of the scan kernels
constexpr int kScanChunk = 4096;   // slices staged in LDS at a time

// b[k]: largest radius of slice lo + k, -1 for an empty slice.  Writes
// dr[lo..hi) and returns rmax after slice hi - 1; one wavefront.
__device__ __forceinline__ double scan_walk(const double *b, int lo, int hi, double rmax, int ngr, double *dr) {
    const int lane = threadIdx.x & 63;
    for (int base = lo; base < hi; base += 64) {
        const int i = base + lane;
        const bool valid = i < hi;
        const double bi = valid ? b[i - lo] : -1.0;
        const bool nonempty = bi >= 0.0;
        double mine = rmax;
        int start = 0;
        for (;;) {
            const unsigned long long m = __ballot(lane >= start && nonempty && bi > rmax);
            if (m == 0) {
                if (lane >= start) mine = rmax;
                break;
            }
            const int jl = __ffsll((long long)m) - 1;
            if (lane >= start && lane < jl) mine = rmax;
            rmax = __shfl(bi, jl) * 1.5;
            if (lane == jl) mine = rmax;
            start = jl + 1;
        }
        if (valid) dr[i] = mine / (double)(ngr - 1);
    }
    return rmax;
}

// The whole scan on one workgroup (any multiple of 64 threads, every thread
// of it calls this); lds: kScanChunk doubles.
__device__ __forceinline__ void scan_staged(int ns, const long long *off, const double *rbound, double *rmax_state,
                                            int ngr, double *dr, double *lds) {
    double rmax = rmax_state[0];
    for (int lo = 0; lo < ns; lo += kScanChunk) {
        const int hi = min(ns, lo + kScanChunk);
        for (int i = lo + (int)threadIdx.x; i < hi; i += (int)blockDim.x)
            lds[i - lo] = (off[i + 1] - off[i] >= 1) ? rbound[i] : -1.0;
        __syncthreads();
        if (threadIdx.x < 64) rmax = scan_walk(lds, lo, hi, rmax, ngr, dr);
        __syncthreads();
    }
    if (threadIdx.x == 0) rmax_state[0] = rmax;
}

}  // namespace g4mi
