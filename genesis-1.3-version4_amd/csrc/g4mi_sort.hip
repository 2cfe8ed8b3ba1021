// Particle migration of one4one beams on the device, Beam::sort
// (src/Core/Beam.cpp:81-94) with Sorting::fillPushVectors / localSort
// (src/Util/Sorting.cpp:74-145, 367-428) for one rank: a particle at
// s = slen a + theta outside [0, slen ns] is dropped, every other one moves
// to slice a + floor(theta / slen) with theta -= slen atar.
//
// Layout: keys = destination slice (ns for a dropped particle) and the
// particle's index go through a stable LSD radix sort (4-bit digits, below),
// so each slice keeps its particles in source order; one gather writes the
// six coordinates in the new CSR order and a binary search per slice gives
// the new offsets.  The reference's own order within a slice (swap with the
// last particle, push_back) differs; only the set of particles in each
// slice is the same.
#include <algorithm>
#include <cstring>

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

// ---- stable LSD radix sort of (key, value) pairs, 4 bits per pass --------
//
// Per pass: (1) each tile of 2048 pairs counts its digits, hist[d nb + b];
// (2) an exclusive scan of hist in that digit-major order gives every
// (digit, tile) its first output position; (3) each tile scatters its pairs
// there, a pair's rank among the tile's pairs of its digit counted in index
// order (thread t owns the 8 consecutive pairs t*8.., so the thread order is
// the index order), which keeps equal keys in input order.

constexpr int kRadixNT = 256, kRadixItems = 8, kRadixTile = kRadixNT * kRadixItems, kRadixBins = 16;

__global__ void __launch_bounds__(kRadixNT) k_radix_hist(const int *__restrict__ keys, long long n, int shift,
                                                         int *__restrict__ hist, int nb) {
    __shared__ int h[kRadixBins];
    const int t = threadIdx.x;
    if (t < kRadixBins) h[t] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kRadixTile;
    for (int k = 0; k < kRadixItems; k++) {
        const long long i = base + (long long)k * kRadixNT + t;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kRadixBins - 1)], 1);
    }
    __syncthreads();
    if (t < kRadixBins) hist[(long long)t * nb + blockIdx.x] = h[t];
}

// exclusive scan of the block's kRadixNT x IT ints, thread t holding entries
// t*IT .. t*IT+IT-1 in v; returns the block total
template <int IT>
__device__ int radix_block_scan(int (&v)[IT], int *lds) {
    const int t = threadIdx.x;
    int run = 0;
    for (int k = 0; k < IT; k++) {
        const int c = v[k];
        v[k] = run;
        run += c;
    }
    lds[t] = run;
    __syncthreads();
    for (int o = 1; o < kRadixNT; o <<= 1) {  // Hillis-Steele inclusive scan of the thread totals
        const int a = t >= o ? lds[t - o] : 0;
        __syncthreads();
        lds[t] += a;
        __syncthreads();
    }
    const int excl = lds[t] - run, total = lds[kRadixNT - 1];
    for (int k = 0; k < IT; k++) v[k] += excl;
    __syncthreads();
    return total;
}

// exclusive scan of m ints in place, chunk totals to sums[]
__global__ void __launch_bounds__(kRadixNT) k_radix_scan_chunks(int *__restrict__ a, long long m,
                                                                int *__restrict__ sums) {
    __shared__ int lds[kRadixNT];
    const long long base = (long long)blockIdx.x * kRadixTile + (long long)threadIdx.x * kRadixItems;
    int v[kRadixItems];
    for (int k = 0; k < kRadixItems; k++) v[k] = base + k < m ? a[base + k] : 0;
    const int total = radix_block_scan(v, lds);
    for (int k = 0; k < kRadixItems; k++)
        if (base + k < m) a[base + k] = v[k];
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// exclusive scan of the chunk totals (one workgroup, any count)
__global__ void __launch_bounds__(kRadixNT) k_radix_scan_sums(int *__restrict__ sums, int nchunks) {
    __shared__ int lds[kRadixNT];
    int carry = 0;
    for (int c0 = 0; c0 < nchunks; c0 += kRadixTile) {
        const int base = c0 + threadIdx.x * kRadixItems;
        int v[kRadixItems];
        for (int k = 0; k < kRadixItems; k++) v[k] = base + k < nchunks ? sums[base + k] : 0;
        const int total = radix_block_scan(v, lds);
        for (int k = 0; k < kRadixItems; k++)
            if (base + k < nchunks) sums[base + k] = v[k] + carry;
        carry += total;
    }
}

__global__ void __launch_bounds__(kRadixNT) k_radix_scan_add(int *__restrict__ a, long long m,
                                                             const int *__restrict__ sums) {
    const long long base = (long long)blockIdx.x * kRadixTile;
    const int add = sums[blockIdx.x];
    for (int k = threadIdx.x; k < kRadixTile; k += kRadixNT)
        if (base + k < m) a[base + k] += add;
}

__global__ void __launch_bounds__(kRadixNT) k_radix_scatter(const int *__restrict__ kin, const int *__restrict__ vin,
                                                            int *__restrict__ kout, int *__restrict__ vout,
                                                            long long n, int shift, const int *__restrict__ pos0,
                                                            int nb) {
    __shared__ int cnt[kRadixBins * kRadixNT];  // [digit][thread], then its exclusive scan
    __shared__ int lds[kRadixNT];
    __shared__ int dstart[kRadixBins];
    const int t = threadIdx.x, b = blockIdx.x;
    const long long base = (long long)b * kRadixTile + (long long)t * kRadixItems;
    int key[kRadixItems], val[kRadixItems];
    for (int k = 0; k < kRadixItems; k++) {
        key[k] = base + k < n ? kin[base + k] : -1;
        val[k] = base + k < n ? vin[base + k] : 0;
    }
    for (int d = 0; d < kRadixBins; d++) cnt[d * kRadixNT + t] = 0;
    for (int k = 0; k < kRadixItems; k++)
        if (key[k] >= 0) cnt[((key[k] >> shift) & (kRadixBins - 1)) * kRadixNT + t] += 1;
    __syncthreads();
    // flattened [digit][thread] exclusive scan: thread t scans entries t*16..t*16+15
    int v[kRadixBins];
    for (int k = 0; k < kRadixBins; k++) v[k] = cnt[t * kRadixBins + k];
    __syncthreads();
    radix_block_scan(v, lds);
    for (int k = 0; k < kRadixBins; k++) cnt[t * kRadixBins + k] = v[k];
    __syncthreads();
    if (t < kRadixBins) dstart[t] = cnt[t * kRadixNT];
    __syncthreads();
    for (int k = 0; k < kRadixItems; k++) {
        if (key[k] < 0) continue;
        const int d = (key[k] >> shift) & (kRadixBins - 1);
        int &c = cnt[d * kRadixNT + t];  // this thread's entry: no other thread touches it
        const long long p = (long long)pos0[(long long)d * nb + b] + (c - dstart[d]);
        c += 1;
        kout[p] = key[k];
        vout[p] = val[k];
    }
}

static int radix_passes(int ns) {  // keys are 0..ns
    int b = 1;
    while ((1LL << b) <= (long long)ns) b++;
    return (b + 3) / 4;
}

int sort_work_bytes(int ns, long long n, size_t *bytes) {
    (void)ns;
    const long long nb = std::max(1LL, (n + kRadixTile - 1) / kRadixTile);
    const long long m = (long long)kRadixBins * nb, nchunks = (m + kRadixTile - 1) / kRadixTile;
    *bytes = (size_t)(m + nchunks) * sizeof(int);
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
    const long long nb = (n + kRadixTile - 1) / kRadixTile;
    const long long m = (long long)kRadixBins * nb, nchunks = (m + kRadixTile - 1) / kRadixTile;
    if ((size_t)(m + nchunks) * sizeof(int) > w.temp_bytes) {
        set_error("sort: work buffer smaller than the particle count needs");
        return G4MI_E_ARG;
    }
    int *hist = (int *)w.temp, *sums = hist + m;
    hipLaunchKernelGGL(k_sort_keys, dim3(ns), dim3(256), 0, st, ns, off, th, slen, 1.0 / slen, w.keys, w.vals,
                       w.th_new);
    int *kin = w.keys, *vin = w.vals, *kout = w.keys_out, *vout = w.vals_out;
    const int passes = radix_passes(ns);
    for (int p = 0; p < passes; p++) {
        const int shift = 4 * p;
        hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nb), dim3(kRadixNT), 0, st, kin, n, shift, hist, (int)nb);
        hipLaunchKernelGGL(k_radix_scan_chunks, dim3((unsigned)nchunks), dim3(kRadixNT), 0, st, hist, m, sums);
        hipLaunchKernelGGL(k_radix_scan_sums, dim3(1), dim3(kRadixNT), 0, st, sums, (int)nchunks);
        hipLaunchKernelGGL(k_radix_scan_add, dim3((unsigned)nchunks), dim3(kRadixNT), 0, st, hist, m, sums);
        hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nb), dim3(kRadixNT), 0, st, kin, vin, kout, vout, n,
                           shift, hist, (int)nb);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // the sorted pairs are in kin/vin after the last swap
    const long long blocks = std::min<long long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_sort_gather, dim3((unsigned)blocks), dim3(256), 0, st, n, vin, w.th_new, g, x, y, px, py,
                       w.g, w.th, w.x, w.y, w.px, w.py);
    hipLaunchKernelGGL(k_sort_bounds, dim3((ns + 256) / 256), dim3(256), 0, st, ns, n, kin, off_new);
    G4_HIP(hipGetLastError());
    return G4MI_OK;
}

}  // namespace g4mi
