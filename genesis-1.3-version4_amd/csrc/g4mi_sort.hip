// Particle migration of one4one beams on the device, Beam::sort
// (src/Core/Beam.cpp:81-94) with Sorting::fillPushVectors / localSort
// (src/Util/Sorting.cpp:74-145, 367-428) for one rank: a particle at
// s = slen a + theta outside [0, slen ns] is dropped, every other one moves
// to slice a + floor(theta / slen) with theta -= slen atar.
//
// Layout: keys = destination slice (ns for a dropped particle) and the
// particle's index go through a stable radix sort (rocPRIM), so each slice
// keeps its particles in source order; one gather writes the six
// coordinates in the new CSR order and a binary search per slice gives the
// new offsets.  The reference's own order within a slice (swap with the
// last particle, push_back) differs; only the set of particles in each
// slice is the same.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "g4mi_sort.h"

namespace g4mi {

__global__ void __launch_bounds__(256) k_sort_keys(int ns, const long long *__restrict__ off,
                                                   const double *__restrict__ th, double slen, double invslen,
                                                   int *__restrict__ keys, int *__restrict__ vals,
                                                   double *__restrict__ th_new) {
    const int a = blockIdx.x;
    const long long b0 = off[a], b1 = off[a + 1];
    const double smax = slen * ns;
    for (long long i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        const double theta = th[i];
        const double s = slen * a + theta;
        int key = ns;
        double tn = theta;
        if (!((s < 0.0) || (s > smax))) {
            const int atar = (int)floor(theta * invslen);
            const int dest = a + atar;
            // s exactly at the upper edge would index one past the last slice,
            // and s = 0 can round to a target one below slice 0: both dropped
            if (dest >= 0 && dest < ns) {
                key = dest;
                tn = theta - slen * (atar);
            }
        }
        keys[i] = key;
        vals[i] = (int)i;
        th_new[i] = tn;
    }
}

__global__ void k_sort_gather(long long n, const int *__restrict__ perm, const double *__restrict__ th_new,
                              const double *__restrict__ g, const double *__restrict__ x,
                              const double *__restrict__ y, const double *__restrict__ px,
                              const double *__restrict__ py, double *__restrict__ og, double *__restrict__ oth,
                              double *__restrict__ ox, double *__restrict__ oy, double *__restrict__ opx,
                              double *__restrict__ opy) {
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n;
         j += (long long)gridDim.x * blockDim.x) {
        const int src = perm[j];
        og[j] = g[src];
        oth[j] = th_new[src];
        ox[j] = x[src];
        oy[j] = y[src];
        opx[j] = px[src];
        opy[j] = py[src];
    }
}

// off[s] = first position with key >= s in the sorted keys, s = 0..ns
__global__ void k_sort_bounds(int ns, long long n, const int *__restrict__ keys, long long *__restrict__ off) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > ns) return;
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (keys[mid] < s) lo = mid + 1; else hi = mid;
    }
    off[s] = lo;
}

static unsigned key_bits(int ns) {  // keys are 0..ns
    unsigned b = 1;
    while ((1LL << b) <= (long long)ns) b++;
    return b;
}

int sort_work_bytes(int ns, long long n, size_t *bytes) {
    size_t tb = 0;
    hipError_t e = rocprim::radix_sort_pairs((void *)nullptr, tb, (const int *)nullptr, (int *)nullptr,
                                             (const int *)nullptr, (int *)nullptr, (unsigned int)n, 0, key_bits(ns));
    if (e != hipSuccess) {
        set_error(std::string("radix sort size query: ") + hipGetErrorString(e));
        return G4MI_E_HIP;
    }
    *bytes = tb;
    return G4MI_OK;
}

int launch_sort_one4one(hipStream_t st, int ns, long long n, const long long *off, double slen, double *g,
                        double *th, double *x, double *y, double *px, double *py, const SortWork &w,
                        long long *off_new) {
    if (n <= 0) {
        hipLaunchKernelGGL(k_sort_bounds, dim3((ns + 256) / 256), dim3(256), 0, st, ns, 0LL, w.keys_out, off_new);
        G4_HIP(hipGetLastError());
        return G4MI_OK;
    }
    if (n > 0x7fffffffLL) {
        set_error("sort: more than 2^31 particles");
        return G4MI_E_ARG;
    }
    hipLaunchKernelGGL(k_sort_keys, dim3(ns), dim3(256), 0, st, ns, off, th, slen, 1.0 / slen, w.keys, w.vals,
                       w.th_new);
    size_t tb = w.temp_bytes;
    G4_HIP(rocprim::radix_sort_pairs(w.temp, tb, w.keys, w.keys_out, w.vals, w.vals_out, (unsigned int)n, 0,
                                     key_bits(ns), st));
    const long long blocks = std::min<long long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_sort_gather, dim3((unsigned)blocks), dim3(256), 0, st, n, w.vals_out, w.th_new, g, x, y,
                       px, py, w.g, w.th, w.x, w.y, w.px, w.py);
    hipLaunchKernelGGL(k_sort_bounds, dim3((ns + 256) / 256), dim3(256), 0, st, ns, n, w.keys_out, off_new);
    G4_HIP(hipGetLastError());
    return G4MI_OK;
}

}  // namespace g4mi
