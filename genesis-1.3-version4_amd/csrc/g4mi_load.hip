// FP64 beam loading on the device: quiet Hammersley loading per slice
// (QuietLoading::loadQuiet, src/Loading/QuietLoading.cpp:50-260) and the
// shot-noise phase kicks (ShotNoise::applyShotNoise,
// src/Loading/ShotNoise.cpp:29-68).
//
// Quiet loading is one workgroup per slice.  Beamlet i of a slice draws
// its six raw coordinates from the radical inverse of q = start + 1 + i in
// the bases 2, 3, 5, 7, 11, 13 (the Hammersley counters of the slice's
// sequence all start at `start`, the host's first ran2 draw of the slice
// seed) and the inverse complementary error function.  The moments the
// reference takes over the slice (exact mean and rms of the energy, x-px
// and y-py decorrelation, unit momentum rms) are block sums here, so the
// raw draws are recomputed in each of the three passes instead of being
// kept: the third pass writes the nbins mirrored macroparticles of every
// beamlet in place, which would overwrite raw draws of later beamlets.
//
// The shot noise draws come from one sequential ran2 stream over all
// slices, which has no skip-ahead; the host draws them in the reference's
// order (phase and amplitude per harmonic and beamlet) and the device adds
// the kicks.
#include <algorithm>

#include <hip/hip_runtime.h>

#include "g4mi.h"
#include "g4mi_load.h"

namespace g4mi {
namespace {

constexpr int kLoadNT = 256;

// radical inverse of q in base b, digit by digit as the host's Halton
// sequence (src/Util/Hammerslay.cpp:41-49); the same IEEE operations
__device__ double radical_inverse(unsigned int q, double b) {
    const unsigned int ib = (unsigned int)b;
    double v = 0, scale = 1.0;
    do {
        scale = scale / b;
        const unsigned int d = q / ib;
        v += (q - b * d) * scale;
        q = d;
    } while (q > 0);
    return v;
}

// inverse complementary error function (T. Ooura's method, the one of
// src/Util/Inverfc.cpp:16-64: a rational first guess refined by one
// high-order correction; Ooura's coefficients)
__device__ double inverfc(double y) {
    const double z = (y > 1) ? 2 - y : y;
    const double w = 0.916461398268964 - log(z);
    double u = sqrt(w);
    double s = (log(u) + 0.488826640273108) / w;
    double t = 1 / (u + 0.231729200323405);
    double x = u * (1 - s * (s * 0.124610454613712 + 0.5)) -
               ((((-0.0728846765585675 * t + 0.269999308670029) * t + 0.150689047360223) * t +
                 0.116065025341614) * t + 0.499999303439796) * t;
    t = 3.97886080735226 / (x + 3.97886080735226);
    u = t - 0.5;
    s = 0.00112648096188977922;
    s = s * u + 1.05739299623423047e-4;
    s = s * u - 0.00351287146129100025;
    s = s * u - 7.71708358954120939e-4;
    s = s * u + 0.00685649426074558612;
    s = s * u + 0.00339721910367775861;
    s = s * u - 0.011274916933250487;
    s = s * u - 0.0118598117047771104;
    s = s * u + 0.0142961988697898018;
    s = s * u + 0.0346494207789099922;
    s = s * u + 0.00220995927012179067;
    s = s * u - 0.0743424357241784861;
    s = s * u - 0.105872177941595488;
    s = s * u + 0.0147297938331485121;
    s = s * u + 0.316847638520135944;
    s = s * u + 0.713657635868730364;
    s = s * u + 1.05375024970847138;
    s = s * u + 1.21448730779995237;
    s = s * u + 1.16374581931560831;
    s = s * u + 0.956464974744799006;
    s = s * u + 0.686265948274097816;
    s = s * u + 0.434397492331430115;
    s = s * u + 0.244044510593190935;
    s = s * t - z * exp(x * x - 0.120782237635245222);
    x = x + s * (x * s + 1);
    return (y > 1) ? -x : x;
}

// deterministic block sum of K doubles (wave butterfly, then waves in order)
template <int K>
__device__ void load_block_sum(double (&v)[K], double *lds) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
    for (int k = 0; k < K; k++)
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
    __syncthreads();
    if (lane == 0)
        for (int k = 0; k < K; k++) lds[w * K + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = lds[k];
        for (int q = 1; q < kLoadNT / 64; q++) s += lds[q * K + k];
        v[k] = s;
    }
    __syncthreads();
}

struct Raw {
    double g, x, y, px, py;
};

__device__ Raw raw_draw(unsigned int q) {
    Raw r;
    r.g = inverfc(2 * radical_inverse(q, 3));
    r.x = inverfc(2 * radical_inverse(q, 5));
    r.y = inverfc(2 * radical_inverse(q, 7));
    r.px = inverfc(2 * radical_inverse(q, 11));
    r.py = inverfc(2 * radical_inverse(q, 13));
    return r;
}

__global__ void __launch_bounds__(kLoadNT) k_quiet_load(const long long *__restrict__ off,
                                                       const unsigned int *__restrict__ starts,
                                                       const g4mi_slice_beam *__restrict__ slices, int nbins,
                                                       double theta0, double *__restrict__ g,
                                                       double *__restrict__ th, double *__restrict__ x,
                                                       double *__restrict__ y, double *__restrict__ px,
                                                       double *__restrict__ py) {
    __shared__ double lds[10 * (kLoadNT / 64)];
    const int is = blockIdx.x, t = threadIdx.x;
    const long long a = off[is];
    const int npart = (int)(off[is + 1] - a);
    const int mpart = npart / nbins;
    const double dtheta = 1. / (double)nbins;
    const double norm = mpart > 0 ? 1. / (double)mpart : 0;
    const g4mi_slice_beam sl = slices[is];
    const unsigned int start = starts[is];
    // pass 1: energy moments and the x-px, y-py sums (QuietLoading.cpp:95-156)
    double s1[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = t; i < mpart; i += kLoadNT) {
        const Raw r = raw_draw(start + 1 + (unsigned int)i);
        s1[0] += r.g;
        s1[1] += r.g * r.g;
        s1[2] += r.x;
        s1[3] += r.x * r.x;
        s1[4] += r.px;
        s1[5] += r.x * r.px;
        s1[6] += r.y;
        s1[7] += r.y * r.y;
        s1[8] += r.py;
        s1[9] += r.y * r.py;
    }
    load_block_sum<10>(s1, lds);
    const double m1 = s1[0] * norm;
    double m2 = sqrt(fabs(s1[1] * norm - m1 * m1));
    if (m2 > 0) m2 = 1 / m2;
    m2 *= sl.delgam;
    // decorrelate(a, b): b -= mean(b) + slope a; a -> (a - mean) / rms
    auto decor = [&](double sa, double saa, double sb, double sab, double &ma, double &ra, double &mb,
                     double &slope) {
        ma = sa * norm;
        mb = sb * norm;
        ra = sqrt(fabs(saa * norm - ma * ma));
        if (ra > 0) ra = 1. / ra;
        slope = (sab * norm - ma * mb) * ra * ra;
    };
    double xa, xr, pxm, xs, ya, yr, pym, ys;
    decor(s1[2], s1[3], s1[4], s1[5], xa, xr, pxm, xs);
    decor(s1[6], s1[7], s1[8], s1[9], ya, yr, pym, ys);
    // pass 2: the momenta's mean and rms after the decorrelation
    double s2[4] = {0, 0, 0, 0};
    for (int i = t; i < mpart; i += kLoadNT) {
        const Raw r = raw_draw(start + 1 + (unsigned int)i);
        double bx = r.px - pxm;
        bx -= xs * r.x;
        double by = r.py - pym;
        by -= ys * r.y;
        s2[0] += bx;
        s2[1] += bx * bx;
        s2[2] += by;
        s2[3] += by * by;
    }
    load_block_sum<4>(s2, lds);
    const double a1 = s2[0] * norm, b1 = s2[2] * norm;
    double a2 = sqrt(fabs(s2[1] * norm - a1 * a1));
    if (a2 > 0) a2 = 1. / a2;
    double b2 = sqrt(fabs(s2[3] * norm - b1 * b1));
    if (b2 > 0) b2 = 1. / b2;
    // Twiss scaling, centroids (QuietLoading.cpp:158-209)
    const double sigx = sqrt(sl.ex * sl.betax / sl.gamma);
    const double sigy = sqrt(sl.ey * sl.betay / sl.gamma);
    const double sigpx = sqrt(sl.ex / sl.betax / sl.gamma);
    const double sigpy = sqrt(sl.ey / sl.betay / sl.gamma);
    const double corx = -sl.alphax / sl.betax;
    const double cory = -sl.alphay / sl.betay;
    // bunching: invert sin(x)/x for the requested value (QuietLoading.cpp:225-245)
    const double pi = 2 * asin(1.);
    double bl = sl.bunch;
    if (bl != 0) {
        double lo = 0, hi = 0.9999, mid = 0;
        while ((hi - lo) > 1e-6) {
            mid = 0.5 * (lo + hi);
            const double probe = sin((1 - mid) * pi) / ((1 - mid) * pi);
            if (bl < probe) hi = mid;
            else lo = mid;
        }
        bl = mid;
    }
    const bool modulate = (bl != 0) || (sl.emod != 0);
    // pass 3: every beamlet's nbins mirrored macroparticles
    for (int i = t; i < mpart; i += kLoadNT) {
        const unsigned int q = start + 1 + (unsigned int)i;
        const Raw r = raw_draw(q);
        const double gi = (r.g - m1) * m2 + sl.gamma;
        double bx = r.px - pxm;
        bx -= xs * r.x;
        double by = r.py - pym;
        by -= ys * r.y;
        double xi = (r.x - xa) * xr, yi = (r.y - ya) * yr;
        double pxi = (bx - a1) * a2, pyi = (by - b1) * b2;
        xi *= sigx;
        yi *= sigy;
        pxi = sigpx * pxi + corx * xi;
        pyi = sigpy * pyi + cory * yi;
        pxi *= gi;
        pyi *= gi;
        xi += sl.xcen;
        yi += sl.ycen;
        pxi += sl.pxcen;
        pyi += sl.pycen;
        const double thi = radical_inverse(q, 2) * dtheta;
        for (int j = 0; j < nbins; j++) {
            const long long k = a + (long long)nbins * i + j;
            double gk = gi, tk = (thi + j * dtheta) * theta0;
            if (modulate) {
                gk -= sl.emod * sin(tk - sl.emodphase);
                if (sl.bunch > 0.1) tk += -bl * tk + bl * pi + sl.bunchphase + pi;
                else tk -= 2 * sl.bunch * sin(tk - sl.bunchphase);
            }
            g[k] = gk;
            th[k] = tk;
            x[k] = xi;
            y[k] = yi;
            px[k] = pxi;
            py[k] = pyi;
        }
    }
}

// shot noise: dth[k] = -sum_h an[h][i1] sin((h + 1) th[k] + phi[h][i1]) for
// particle k of beamlet i1, all harmonics from the phases before the kicks
// (ShotNoise.cpp:48-66); draws[(slice base + h mpart + i1) * 2] = {phi, an}
__global__ void __launch_bounds__(256) k_shot_noise(const long long *__restrict__ off,
                                                    const long long *__restrict__ dbase, int nbins, int nh,
                                                    const double *__restrict__ draws, double *__restrict__ th) {
    const int is = blockIdx.y;
    const long long a = off[is];
    const int npart = (int)(off[is + 1] - a);
    const int mpart = npart / nbins;
    const long long db = dbase[is];
    if (db < 0) return;  // no shot noise for this slice
    for (int k = blockIdx.x * 256 + threadIdx.x; k < npart; k += gridDim.x * 256) {
        const int i1 = k / nbins;
        const double t0 = th[a + k];
        double d = 0;
        for (int h = 0; h < nh; h++) {
            const double *p = draws + 2 * (db + (long long)h * mpart + i1);
            d -= p[1] * sin(t0 * (h + 1) + p[0]);
        }
        th[a + k] = t0 + d;
    }
}

// mixed form: gamma and theta as FP32 offsets on the reference energy and
// the slice's phase base, the rest rounded to FP32 (as the host upload)
__global__ void __launch_bounds__(256) k_slice_theta_base(const long long *__restrict__ off,
                                                          const double *__restrict__ th,
                                                          double *__restrict__ base) {
    __shared__ double lds[4];
    const int is = blockIdx.x, t = threadIdx.x;
    const long long a = off[is], b = off[is + 1];
    double v[1] = {0};
    for (long long i = a + t; i < b; i += 256) v[0] += th[i];
    const int lane = t & 63, w = t >> 6;
    for (int o = 32; o > 0; o >>= 1) v[0] += __shfl_xor(v[0], o);
    if (lane == 0) lds[w] = v[0];
    __syncthreads();
    if (t == 0) {
        const double s = lds[0] + lds[1] + lds[2] + lds[3];
        const double twopi = 4. * asin(1.0);
        base[is] = b > a ? twopi * floor(s / (double)(b - a) / twopi) : 0.0;
    }
}

__global__ void __launch_bounds__(256) k_sum_gamma(const double *__restrict__ g, long long n,
                                                   double *__restrict__ out) {
    __shared__ double lds[4];
    const int t = threadIdx.x;
    double v = 0;
    for (long long i = (long long)blockIdx.x * 256 + t; i < n; i += (long long)gridDim.x * 256) v += g[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((t & 63) == 0) lds[t >> 6] = v;
    __syncthreads();
    if (t == 0) out[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}

__global__ void __launch_bounds__(256) k_to_offsets(const long long *__restrict__ off, int ns, long long n,
                                                    const double *__restrict__ base, double gref,
                                                    const double *__restrict__ g, const double *__restrict__ th,
                                                    const double *__restrict__ x, const double *__restrict__ y,
                                                    const double *__restrict__ px, const double *__restrict__ py,
                                                    float *__restrict__ fg, float *__restrict__ ft,
                                                    float *__restrict__ fx, float *__restrict__ fy,
                                                    float *__restrict__ fpx, float *__restrict__ fpy) {
    const int is = blockIdx.y;
    const long long a = off[is], b = off[is + 1];
    const double bs = base[is];
    for (long long i = a + blockIdx.x * 256 + threadIdx.x; i < b; i += (long long)gridDim.x * 256) {
        fg[i] = (float)(g[i] - gref);
        ft[i] = (float)(th[i] - bs);
        fx[i] = (float)x[i];
        fy[i] = (float)y[i];
        fpx[i] = (float)px[i];
        fpy[i] = (float)py[i];
    }
}

}  // namespace

void launch_quiet_load(hipStream_t s, int ns, const long long *off, const unsigned int *starts,
                       const g4mi_slice_beam *slices, int nbins, double theta0, double *g, double *th, double *x,
                       double *y, double *px, double *py) {
    hipLaunchKernelGGL(k_quiet_load, dim3(ns), dim3(kLoadNT), 0, s, off, starts, slices, nbins, theta0, g, th, x, y,
                       px, py);
}

void launch_shot_noise(hipStream_t s, int ns, const long long *off, const long long *dbase, int nbins, int nh,
                       const double *draws, double *th, int maxpart) {
    const unsigned gx = (unsigned)std::max(1, std::min(64, (maxpart + 255) / 256));
    hipLaunchKernelGGL(k_shot_noise, dim3(gx, (unsigned)ns), dim3(256), 0, s, off, dbase, nbins, nh, draws, th);
}

void launch_theta_base(hipStream_t s, int ns, const long long *off, const double *th, double *base) {
    hipLaunchKernelGGL(k_slice_theta_base, dim3(ns), dim3(256), 0, s, off, th, base);
}

void launch_sum_gamma(hipStream_t s, const double *g, long long n, double *out, int nblocks) {
    hipLaunchKernelGGL(k_sum_gamma, dim3(nblocks), dim3(256), 0, s, g, n, out);
}

void launch_to_offsets(hipStream_t s, int ns, const long long *off, long long n, const double *base, double gref,
                       const double *g, const double *th, const double *x, const double *y, const double *px,
                       const double *py, float *fg, float *ft, float *fx, float *fy, float *fpx, float *fpy,
                       int maxpart) {
    const unsigned gx = (unsigned)std::max(1, std::min(64, (maxpart + 255) / 256));
    hipLaunchKernelGGL(k_to_offsets, dim3(gx, (unsigned)ns), dim3(256), 0, s, off, ns, n, base, gref, g, th, x, y,
                       px, py, fg, ft, fx, fy, fpx, fpy);
}

}  // namespace g4mi
