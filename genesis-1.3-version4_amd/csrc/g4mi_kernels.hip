// FP64 device kernels of the undulator step (plain form).
//
// This translation unit is compiled with -ffp-contract=off so every FP64
// expression rounds exactly as the reference's (non-FMA) CPU build: with
// identical inputs, + - * / sqrt agree bit for bit; sin/cos/log/exp/hypot
// may differ from glibc by an ulp.  Each kernel cites the reference loop it
// restates; expressions keep the reference's association order.
#include "g4mi_kernels.h"
#include "g4mi_scan.h"

#include <algorithm>

namespace g4mi {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    // GCC's expansion of std::complex<double> * std::complex<double>
    // (finite operands): (ar br - ai bi, ar bi + ai br).
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
    return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
    return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) {
    return make_double2(a.x * s, a.y * s);
}
// Smith's 1962 complex division, the algorithm of libgcc's __divdc3 for
// finite operands (the host's std::complex<double> operator/).
__device__ __forceinline__ double2 cdiv(double2 n, double2 d) {
    double a = n.x, b = n.y, c = d.x, e = d.y;
    if (fabs(c) < fabs(e)) {
        double ratio = c / e;
        double denom = (c * ratio) + e;
        return make_double2(((a * ratio) + b) / denom, ((b * ratio) - a) / denom);
    }
    double ratio = e / c;
    double denom = (e * ratio) + c;
    return make_double2(((b * ratio) + a) / denom, (b - (a * ratio)) / denom);
}
// libstdc++ pow(complex<double>, int): repeated squaring, reciprocal for
// negative exponents.
__device__ __forceinline__ double2 cpow_int(double2 z, int n) {
    unsigned un = n < 0 ? -(unsigned)n : (unsigned)n;
    double2 x = z;
    double2 y = (un % 2) ? x : make_double2(1.0, 0.0);
    while (un >>= 1) {
        x = cmul(x, x);
        if (un % 2) y = cmul(y, x);
    }
    if (n < 0) return cdiv(make_double2(1.0, 0.0), y);
    return y;
}

// Deterministic block reduction of K doubles (fixed tree order).
template <int NT, int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double *lds) {
    const int t = threadIdx.x;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) lds[k * NT + t] = v[k];
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int k = 0; k < K; k++) lds[k * NT + t] += lds[k * NT + t + s];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = lds[k * NT];
    __syncthreads();
}

// ---------------------------------------------------------------------------
// K1: transverse half step.  TrackBeam::track (src/Core/TrackBeam.cpp:8-81)
// with applyDrift/FQuad/DQuad (:84-116), corrector (:118-128) and the
// chicane matrix application (:223-236).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void line_map(int kind, double delz, double q, double &x, double &px,
                                         double gz, double d) {
    if (kind == 0) {
        x += px * delz / gz;
    } else if (kind > 0) {
        double foc = sqrt(q / gz), omg = foc * delz;
        double s, c;
        sincos(omg, &s, &c);
        double a1 = c, a2 = s / foc, a3 = -a2 * foc * foc;
        double xtmp = x - d;
        x = a1 * xtmp + a2 * px / gz + d;
        px = a3 * xtmp * gz + a1 * px;
    } else {
        double foc = sqrt(-q / gz), omg = foc * delz;
        double a1 = cosh(omg), a2 = sinh(omg) / foc, a3 = a2 * foc * foc;
        double xtmp = x - d;
        x = a1 * xtmp + a2 * px / gz + d;
        px = a3 * xtmp * gz + a1 * px;
    }
}

__device__ __forceinline__ void transverse_one(const TransParams &p, double gi, double &xi, double &yi,
                                               double &pxi, double &pyi) {
    if (p.corrector) {
        pxi += p.ccx;
        pyi += p.ccy;
    }
    if (p.chicane) {
        double gz = sqrt(gi * gi - 1 - pxi * pxi - pyi * pyi);
        double t = xi;
        xi = p.m00 * t + p.m01 * pxi / gz;
        pxi = p.m10 * t * gz + p.m11 * pxi;
        t = yi;
        yi = p.m22 * t + p.m23 * pyi / gz;
        pyi = p.m32 * t * gz + p.m33 * pyi;
    }
    double gammaz = sqrt(gi * gi - 1 - p.aw * p.aw - pxi * pxi - pyi * pyi);
    line_map(p.kindx, p.delz, p.qx, xi, pxi, gammaz, p.xoff);
    line_map(p.kindy, p.delz, p.qy, yi, pyi, gammaz, p.yoff);
}

__global__ void __launch_bounds__(256) k_transverse(long long n, const double *__restrict__ g,
                                                    double *__restrict__ x, double *__restrict__ y,
                                                    double *__restrict__ px,
                                                    double *__restrict__ py, TransParams p) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double xi = x[i], yi = y[i], pxi = px[i], pyi = py[i];
    transverse_one(p, g[i], xi, yi, pxi, pyi);
    x[i] = xi;
    y[i] = yi;
    px[i] = pxi;
    py[i] = pyi;
}

// R56 of a chicane, TrackBeam::applyR56 (:260-277).
__global__ void __launch_bounds__(256) k_r56(long long n, const double *__restrict__ g,
                                             double *__restrict__ th, double r56, double gamma0) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    th[i] += r56 * (g[i] - gamma0);
}

// ---------------------------------------------------------------------------
// Per-slice first and second moments of x and y (Beam::getSize,
// src/Core/Beam.cpp:42-67; centroid of EFieldSolver::analyseBeam :137-145).
// mom[4*is + {0,1,2,3}] = sum x, sum x^2, sum y, sum y^2.
// ---------------------------------------------------------------------------
// Fixed-point accumulation in LDS.  A value v goes in as the integer
// round(v s) with s a power of two, added with 64-bit integer atomics:
// integer addition is associative, so a sum does not depend on the order in
// which the lanes' atomics land, and the FP64 step is bit-reproducible run
// to run.  s is the largest power of two with bound * s <= 2^61, where
// bound caps the magnitude of any partial sum; the sum is rounded once,
// when it is read back as sum / s.  The quantum is bound * 2^-61, below the
// FP64 rounding of a sum of that size.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double fx_scale(double bound) {
    if (!(bound > 0) || !(bound < 1e300)) return 1.0;
    int e;
    (void)frexp(bound, &e);  // bound < 2^e
    return ldexp(1.0, 61 - e);
}
__device__ __forceinline__ long long fx(double v, double s) { return __double2ll_rn(v * s); }
__device__ __forceinline__ void fx_add(long long *cell, long long v) {
    atomicAdd((unsigned long long *)cell, (unsigned long long)v);
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_slice_moments(const long long *__restrict__ off,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       double *__restrict__ mom) {
    __shared__ double lds[4 * 256];
    int is = blockIdx.x;
    long long a = off[is], b = off[is + 1];
    double v[4] = {0, 0, 0, 0};
    for (long long i = a + threadIdx.x; i < b; i += 256) {
        double xi = x[i], yi = y[i];
        v[0] += xi;
        v[1] += xi * xi;
        v[2] += yi;
        v[3] += yi * yi;
    }
    block_sum<256, 4>(v, lds);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; k++) mom[4 * is + k] = v[k];
    }
}

// Slice rms-size product for the long-range solver and the centroid of the
// short-range solver.  fsize as Beam::getSize + the <=0 guard of
// EFieldSolver::longRange (:87-95).
__global__ void k_slice_sizes(int ns, const long long *__restrict__ off,
                              const double *__restrict__ current, const double *__restrict__ mom,
                              double *__restrict__ fsize, double *__restrict__ cen) {
    int is = blockIdx.x * blockDim.x + threadIdx.x;
    if (is >= ns) return;
    long long cnt = off[is + 1] - off[is];
    double sz;
    if (current[is] <= 0 || cnt == 0) {
        sz = 1;
    } else {
        double c = (double)cnt;
        double x1 = mom[4 * is] / c, x2 = mom[4 * is + 1] / c;
        double y1 = mom[4 * is + 2] / c, y2 = mom[4 * is + 3] / c;
        sz = sqrt(fabs(x2 - x1 * x1)) * sqrt(fabs(y2 - y1 * y1));
    }
    if (sz <= 0) sz = 1;
    fsize[is] = sz;
    if (cnt > 0) {
        cen[2 * is] = mom[4 * is] / (double)cnt;
        cen[2 * is + 1] = mom[4 * is + 2] / (double)cnt;
    } else {
        cen[2 * is] = 0;
        cen[2 * is + 1] = 0;
    }
}

// K3a: long-range space charge, EFieldSolver::longRange (:103-116) on one
// device: E_i = sum_j sgn(i-j) (1 - sqrt(ds^2/(ds^2+size_j))) scl I_j/size_j.
// One wavefront per slice i: lanes take j = lane, lane+64, ... and the 64
// partial sums are combined in a fixed butterfly order (deterministic).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__global__ void __launch_bounds__(64) k_lsc_long(int ns, const double *__restrict__ cur,
                                                 const double *__restrict__ fsize, double sl,
                                                 double gamma, double scl,
                                                 double *__restrict__ esc) {
    const int i = blockIdx.x, lane = threadIdx.x;
    double e = 0;
    for (int j = lane; j < ns; j += 64) {
        const double sz = fsize[j];
        const double ds = (double)(j - i) * sl * gamma;
        const double coef = 1 - sqrt(ds * ds / (ds * ds + sz));
        const double sgn = (j > i) ? -1. : ((j < i) ? 1. : 0.);
        e += sgn * coef * scl * cur[j] / sz;
    }
    e = wave_sum(e);
    if (lane == 0) esc[i] = e;
}

// Largest radius from the slice centroid (EFieldSolver::analyseBeam
// :150-157); a max is order independent, so this is exact.
__global__ void __launch_bounds__(256) k_sc_rbound(const long long *__restrict__ off,
                                                   const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   const double *__restrict__ cen,
                                                   double *__restrict__ rbound) {
    __shared__ double lds[256];
    int is = blockIdx.x;
    long long a = off[is], b = off[is + 1];
    double xc = cen[2 * is], yc = cen[2 * is + 1];
    double m = 0;
    for (long long i = a + threadIdx.x; i < b; i += 256) {
        double tx = x[i] - xc, ty = y[i] - yc;
        double r2 = tx * tx + ty * ty;
        if (r2 > m) m = r2;
    }
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) rbound[is] = sqrt(lds[0]);
}

// The growing radial grid of the short-range solver (the walk of
// g4mi_scan.h), one workgroup.
__global__ void __launch_bounds__(kScanNT) k_sc_scan(int ns, const long long *__restrict__ off,
                                                     const double *__restrict__ rbound,
                                                     double *__restrict__ rmax_state, int ngr,
                                                     double *__restrict__ dr) {
    __shared__ double lds[kScanChunk];
    scan_staged(ns, off, rbound, rmax_state, ngr, dr, lds);
}

// K3b: short-range space charge of one slice per workgroup,
// EFieldSolver::shortRange (:197-258) with constructLaplaceOperator
// (:176-190) and tridiag (:260-272).  Radial charge histograms are private
// per wavefront and accumulate in fixed point (fx_add), so the result is
// bit-reproducible.
constexpr int kScMaxNg = 256;
__global__ void __launch_bounds__(256) k_sc_solve(
    const long long *__restrict__ off, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ th, const double *__restrict__ cur, const double *__restrict__ cen,
    const double *__restrict__ dr_arr, int ngr, int nz, int nphi, double ks, double gz2,
    double *__restrict__ ez, double *__restrict__ efield) {
    __shared__ long long hist[4][kScMaxNg][2];
    __shared__ double2 celm[kScMaxNg], gam[kScMaxNg], rhs[kScMaxNg];
    __shared__ double vol[kScMaxNg], ldig[kScMaxNg + 1], rlog[kScMaxNg];
    __shared__ double cupp[kScMaxNg], clow[kScMaxNg], cmid[kScMaxNg];
    const int is = blockIdx.x, t = threadIdx.x, w = t >> 6;
    const long long a = off[is], b = off[is + 1];
    const int npart = (int)(b - a);
    for (long long i = a + t; i < b; i += 256) ez[i] = 0;
    if (t == 0) efield[is] = 0;
    if (npart < 1) return;
    const double dr = dr_arr[is], xc = cen[2 * is], yc = cen[2 * is + 1];
    const double pi = 2 * asin(1.0);
    // radial grid tables, one entry per lane (constructLaplaceOperator :176-190)
    for (int j = t; j <= ngr; j += 256) {
        if (j == ngr) {
            ldig[ngr] = 0;
        } else if (j == 0) {
            vol[0] = pi * dr * dr;
            rlog[0] = 0.5;
            ldig[0] = 0;
        } else {
            vol[j] = pi * dr * dr * (2 * j + 1);
            ldig[j] = 2 * pi * j;
            rlog[j] = log((double)(j + 1) / (double)j);
        }
    }
    const double coef = -gz2 / ks / ks;
    const double econst = kVacimp / kEev * cur[is] / ks / (double)npart;
    const double hs = fx_scale((double)npart);  // |cos|, |sin| <= 1: a bin sums at most npart of them
    for (int m = -nphi; m <= nphi; m++) {
        for (int l = 1; l <= nz; l++) {
            __syncthreads();
            for (int k = t; k < 4 * kScMaxNg; k += 256) hist[k / kScMaxNg][k % kScMaxNg][0] = hist[k / kScMaxNg][k % kScMaxNg][1] = 0;
            __syncthreads();
            for (long long i = a + t; i < b; i += 256) {
                double tx = x[i] - xc, ty = y[i] - yc;
                double radi = sqrt(tx * tx + ty * ty);
                int bin = (int)floor(radi / dr);
                bin = (bin >= 0 && bin < ngr) ? bin : (bin >= ngr ? ngr - 1 : 0);
                double s, c;
                sincos(l * th[i], &s, &c);
                double2 v = make_double2(c, -s);
                if (m != 0)  // cwork^0 is exactly 1: the azimuthal factor only for m != 0
                    v = cmul(cpow_int(make_double2(tx / radi, ty / radi), -m), v);
                fx_add(&hist[w][bin][0], fx(v.x, hs));
                fx_add(&hist[w][bin][1], fx(v.y, hs));
            }
            __syncthreads();
            // operator coefficients (:240-245) and source of every radial cell,
            // one cell per lane; the sequential recurrence then only reads them
            for (int i = t; i < ngr; i += 256) {
                double lm = -ldig[i] - ldig[i + 1] - 2. * pi * m * m * rlog[i];
                if (i == ngr - 1) lm -= 2 * pi * ngr;
                cmid[i] = 1 + coef * lm / l / l / vol[i];
                if (i > 0) {
                    cupp[i] = coef * ldig[i] / l / l / vol[i - 1];
                    clow[i] = coef * ldig[i] / l / l / vol[i];
                }
                const double2 cs = make_double2(
                    (double)(hist[0][i][0] + hist[1][i][0] + hist[2][i][0] + hist[3][i][0]) / hs,
                    (double)(hist[0][i][1] + hist[1][i][1] + hist[2][i][1] + hist[3][i][1]) / hs);
                rhs[i] = cmul(cs, make_double2(0, econst / l / vol[i]));
            }
            __syncthreads();
            if (t == 0) {  // tridiag (:260-272)
                // the operator's entries are real (complex with zero
                // imaginary parts in the reference), so bet and gam are real:
                // one reciprocal per cell on the sequential chain
                double bet = cmid[0], inv = 1.0 / bet;
                celm[0] = cscale(rhs[0], inv);
                for (int i = 1; i < ngr; i++) {
                    const double g = cupp[i] * inv;
                    gam[i].x = g;
                    bet = cmid[i] - clow[i] * g;
                    inv = 1.0 / bet;
                    celm[i] = cscale(csub(rhs[i], cscale(celm[i - 1], clow[i])), inv);
                }
                for (int i = ngr - 2; i > -1; i--) celm[i] = csub(celm[i], cscale(celm[i + 1], gam[i + 1].x));
                if (l == 1) efield[is] = hypot(celm[0].x, celm[0].y);
            }
            __syncthreads();
            for (long long i = a + t; i < b; i += 256) {
                double tx = x[i] - xc, ty = y[i] - yc;
                double radi = sqrt(tx * tx + ty * ty);
                int bin = (int)floor(radi / dr);
                bin = (bin >= 0 && bin < ngr) ? bin : (bin >= ngr ? ngr - 1 : 0);
                double s, c;
                sincos(l * th[i], &s, &c);
                double2 ct = make_double2(c, s);
                if (m != 0) ct = cmul(cpow_int(make_double2(tx / radi, ty / radi), m), ct);
                ez[i] += 2 * cmul(ct, celm[bin]).x;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// K2: longitudinal push.  BeamSolver::advance (src/Core/BeamSolver.cpp:52-85),
// RungeKutta (:89-141), ODE (:144-163), field gather in the order of
// Field::getLLGridpoint (src/Core/Field.cpp:160-174) + :66-79.
// ---------------------------------------------------------------------------
template <int NH>
struct Ode {
    double xks, xku, btpar, ez;
    double2 rpart;
    double2 rph[NH > 0 ? NH : 1];
    double hh[NH > 0 ? NH : 1];
    double k2gg, k2pp;
    __device__ __forceinline__ void eval(double tgam, double tthet) {
        double ztemp1 = -2. / xks;
        double s, c;
        sincos(1.0 * tthet, &s, &c);
        double2 ctmp = cmul(rpart, make_double2(c, -s));
#pragma unroll
        for (int k = 0; k < NH; k++) {  // the harmonic fields, in field order
            double sh, ch;
            sincos(hh[k] * tthet, &sh, &ch);
            ctmp = cadd(ctmp, cmul(rph[k], make_double2(ch, -sh)));
        }
        double btper0 = btpar + ztemp1 * ctmp.x;
        double btpar0 = sqrt(1. - btper0 / (tgam * tgam));
        k2pp += xks * (1. - 1. / btpar0) + xku;
        k2gg += ctmp.y / btpar0 / tgam - ez;
    }
};

template <int NH>
__global__ void __launch_bounds__(256) k_push(const int *__restrict__ blk_slice,
                                              const long long *__restrict__ blk_begin,
                                              const long long *__restrict__ off,
                                              double *__restrict__ g, double *__restrict__ th,
                                              const double *__restrict__ x,
                                              const double *__restrict__ y,
                                              const double *__restrict__ px,
                                              const double *__restrict__ py,
                                              const double2 *__restrict__ field,
                                              const double *__restrict__ esc,
                                              const double *__restrict__ ezsc, PushParams p) {
    const int is = blk_slice[blockIdx.x];
    const long long i = blk_begin[blockIdx.x] + threadIdx.x;
    if (i >= off[is + 1]) return;
    double gamma = g[i];
    double theta = th[i] + p.autophase;
    const double xi = x[i], yi = y[i], pxi = px[i], pyi = py[i];
    double dx = xi - p.ax, dy = yi - p.ay;
    double awloc = (1 + 0.5 * (p.kx * dx * dx + p.ky * dy * dy) + p.gradx * dx + p.grady * dy);
    Ode<NH> o;
    o.xks = p.xks;
    o.xku = p.xku;
    o.btpar = 1 + pxi * pxi + pyi * pyi + p.aw * p.aw * awloc * awloc;
    double eloss = -esc[is] / kEev;
    o.ez = (p.use_sc ? ezsc[i] : 0.0) + eloss;
    if ((xi > -p.gridmax) && (xi < p.gridmax) && (yi > -p.gridmax) && (yi < p.gridmax)) {
        double wx = (xi + p.gridmax) / p.dgrid;
        double wy = (yi + p.gridmax) / p.dgrid;
        int ix = (int)floor(wx), iy = (int)floor(wy);
        wx = 1 + floor(wx) - wx;
        wy = 1 + floor(wy) - wy;
        // (x+gridmax)/dgrid can round up to ngrid-1 within an ulp of the
        // edge; keep the 2x2 stencil inside the slice.
        ix = min(ix, p.ngrid - 2);
        iy = min(iy, p.ngrid - 2);
        int idx = ix + iy * p.ngrid;
        const int islice = (is + p.first) % p.nslice;
        const double2 *f = field + (long long)islice * p.ngrid * p.ngrid;
        double2 f0 = f[idx], f1 = f[idx + 1], f2 = f[idx + p.ngrid], f3 = f[idx + p.ngrid + 1];
        double2 cpart = cscale(cscale(f0, wx), wy);
        cpart = cadd(cpart, cscale(cscale(f1, 1 - wx), wy));
        cpart = cadd(cpart, cscale(cscale(f2, wx), 1 - wy));
        cpart = cadd(cpart, cscale(cscale(f3, 1 - wx), 1 - wy));
        double r = p.rtmp * awloc;
        o.rpart = make_double2(r * cpart.x, r * -cpart.y);
    } else {
        o.rpart = make_double2(0, 0);
    }
#pragma unroll
    for (int k = 0; k < NH; k++) {  // BeamSolver::advance :63-79 for each harmonic field
        const HarmGather &h = p.hg[k];
        o.hh[k] = h.harm;
        o.rph[k] = make_double2(0, 0);
        if ((xi > -h.gridmax) && (xi < h.gridmax) && (yi > -h.gridmax) && (yi < h.gridmax)) {
            double wx = (xi + h.gridmax) / h.dgrid;
            double wy = (yi + h.gridmax) / h.dgrid;
            int ix = min((int)floor(wx), h.ngrid - 2), iy = min((int)floor(wy), h.ngrid - 2);
            wx = 1 + floor(wx) - wx;
            wy = 1 + floor(wy) - wy;
            const int idx = ix + iy * h.ngrid;
            const double2 *f = h.field + (long long)((is + p.first) % p.nslice) * h.ngrid * h.ngrid;
            double2 cpart = cscale(cscale(f[idx], wx), wy);
            cpart = cadd(cpart, cscale(cscale(f[idx + 1], 1 - wx), wy));
            cpart = cadd(cpart, cscale(cscale(f[idx + h.ngrid], wx), 1 - wy));
            cpart = cadd(cpart, cscale(cscale(f[idx + h.ngrid + 1], 1 - wx), 1 - wy));
            const double r = h.rtmp * awloc;
            o.rph[k] = make_double2(r * cpart.x, r * -cpart.y);
        }
    }
    const double delz = p.delz;
    // RungeKutta, :93-141
    double k3gg, k3pp;
    o.k2gg = 0;
    o.k2pp = 0;
    o.eval(gamma, theta);
    double stpz = 0.5 * delz;
    gamma += stpz * o.k2gg;
    theta += stpz * o.k2pp;
    k3gg = o.k2gg;
    k3pp = o.k2pp;
    o.k2gg = 0;
    o.k2pp = 0;
    o.eval(gamma, theta);
    gamma += stpz * (o.k2gg - k3gg);
    theta += stpz * (o.k2pp - k3pp);
    k3gg /= 6;
    k3pp /= 6;
    o.k2gg *= -0.5;
    o.k2pp *= -0.5;
    o.eval(gamma, theta);
    stpz = delz;
    gamma += stpz * o.k2gg;
    theta += stpz * o.k2pp;
    k3gg -= o.k2gg;
    k3pp -= o.k2pp;
    o.k2gg *= 2;
    o.k2pp *= 2;
    o.eval(gamma, theta);
    gamma += stpz * (k3gg + o.k2gg / 6.0);
    theta += stpz * (k3pp + o.k2pp / 6.0);
    g[i] = gamma;
    th[i] = theta;
}

// ---------------------------------------------------------------------------
// K4: source deposition of one beam slice per workgroup,
// FieldSolverADI::advance (src/Core/FieldSolverADI.cpp:13-51).  The slice's
// footprint (bounding box of its in-grid cells + 1) is accumulated in an LDS
// tile in fixed point (fx_add: bit-reproducible), in row bands when it
// exceeds the tile, and written to the slice's source plane.  box[is] = {ix0, iy0, ix1, iy1}
// (inclusive) tells the ADI where the source is non-zero; the plane outside
// the box is never read, so it is never cleared.
// ---------------------------------------------------------------------------
constexpr int kDepTile = 8192;  // complex FP64 cells: 128 KiB of LDS
__global__ void __launch_bounds__(1024) k_deposit(const long long *__restrict__ off,
                                                  const double *__restrict__ g,
                                                  const double *__restrict__ th,
                                                  const double *__restrict__ x,
                                                  const double *__restrict__ y,
                                                  const double *__restrict__ cur,
                                                  double2 *__restrict__ src,
                                                  int4 *__restrict__ box, DepositParams p) {
    extern __shared__ double2 tile_mem[];
    long long *tile = (long long *)tile_mem;  // [cells][2] fixed-point sums
    __shared__ int red[4][1024];
    __shared__ double redd[1024];
    const int is = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
    const long long a = off[is], b = off[is + 1];
    const int ng = p.ngrid;
    int bx0 = 1 << 30, by0 = 1 << 30, bx1 = -1, by1 = -1;
    double pmax = 0;  // largest sqrt(faw2) / gamma of an in-grid particle
    if (p.enabled) {
        for (long long i = a + t; i < b; i += nt) {
            double xi = x[i], yi = y[i];
            if ((xi > -p.gridmax) && (xi < p.gridmax) && (yi > -p.gridmax) && (yi < p.gridmax)) {
                int ix = min((int)floor((xi + p.gridmax) / p.dgrid), ng - 2);
                int iy = min((int)floor((yi + p.gridmax) / p.dgrid), ng - 2);
                bx0 = min(bx0, ix);
                bx1 = max(bx1, ix + 1);
                by0 = min(by0, iy);
                by1 = max(by1, iy + 1);
                const double ddx = xi - p.ax, ddy = yi - p.ay;
                const double faw2 = (1 + p.kx * ddx * ddx + p.ky * ddy * ddy + 2 * (p.gradx * ddx + p.grady * ddy));
                pmax = fmax(pmax, sqrt(faw2) / fabs(g[i]));
            }
        }
    }
    red[0][t] = bx0;
    red[1][t] = by0;
    red[2][t] = -bx1;
    red[3][t] = -by1;
    redd[t] = pmax;
    __syncthreads();
    for (int s = nt / 2; s > 0; s >>= 1) {
        if (t < s) {
            for (int k = 0; k < 4; k++) red[k][t] = min(red[k][t], red[k][t + s]);
            redd[t] = fmax(redd[t], redd[t + s]);
        }
        __syncthreads();
    }
    bx0 = red[0][0];
    by0 = red[1][0];
    bx1 = -red[2][0];
    by1 = -red[3][0];
    pmax = redd[0];
    __syncthreads();
    if (bx1 < 0) {  // nothing deposited
        if (t == 0) box[is] = make_int4(1, 1, 0, 0);
        return;
    }
    bx1 = min(bx1, ng - 1);
    by1 = min(by1, ng - 1);
    if (t == 0) box[is] = make_int4(bx0, by0, bx1, by1);
    const int ww = bx1 - bx0 + 1;
    const int rows_per_band = max(1, kDepTile / ww);
    const double scl0 = p.fcoup * kVacimp * cur[is] * p.xks * p.delz;
    const double scl = scl0 / (4 * kEev * (double)(b - a) * p.dgrid * p.dgrid);
    // no cell sum exceeds sum_i |part_i| <= n max|part|
    const double fs = fx_scale((double)(b - a) * pmax * fabs(scl));
    double2 *plane = src + (long long)is * ng * ng;
    for (int r0 = by0; r0 <= by1; r0 += rows_per_band) {
        const int r1 = min(by1, r0 + rows_per_band - 1);
        const int ncell = (r1 - r0 + 1) * ww;
        for (int k = t; k < 2 * ncell; k += nt) tile[k] = 0;
        __syncthreads();
        // The trip count is the same for every lane (inactive lanes carry
        // zeros), so the lanes of a group of four can combine their values.
        // The quiet loader puts a beamlet's particles in consecutive slots at
        // one transverse position: when the four lanes of a group deposit
        // into the same cells, their sums go in with one set of atomics
        // instead of four (fewer LDS atomic collisions).
        for (long long base = a; base < b; base += nt) {
            const long long i = base + t;
            bool act = i < b;
            double xi = 0, yi = 0;
            if (act) {
                xi = x[i];
                yi = y[i];
                act = (xi > -p.gridmax) && (xi < p.gridmax) && (yi > -p.gridmax) && (yi < p.gridmax);
            }
            int ix = 0, iy = 0;
            double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // (re, im) for cells (ix,iy) (ix+1,iy) (ix,iy+1) (ix+1,iy+1)
            if (act) {
                double wx = (xi + p.gridmax) / p.dgrid;
                double wy = (yi + p.gridmax) / p.dgrid;
                ix = min((int)floor(wx), ng - 2);
                iy = min((int)floor(wy), ng - 2);
                act = !(iy + 1 < r0 || iy > r1);
                if (act) {
                    wx = 1 + floor(wx) - wx;
                    wy = 1 + floor(wy) - wy;
                    double ddx = xi - p.ax, ddy = yi - p.ay;
                    double faw2 = (1 + p.kx * ddx * ddx + p.ky * ddy * ddy + 2 * (p.gradx * ddx + p.grady * ddy));
                    double part = sqrt(faw2) * scl / g[i];
                    double s, c;
                    sincos(p.harm * th[i], &s, &c);
                    const double2 cp = make_double2(s * part, c * part);
                    const double wgt[4] = {wx * wy, (1 - wx) * wy, wx * (1 - wy), (1 - wx) * (1 - wy)};
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        v[2 * k] = wgt[k] * cp.x;
                        v[2 * k + 1] = wgt[k] * cp.y;
                    }
                }
            }
            // all four lanes of the group active and in the same cell?  Every
            // lane takes all three shuffles before any compare: a shuffle
            // from a lane that skipped it would read a stale register.
            const int key = act ? iy * ng + ix : -1 - (t & 3);
            const int k1 = __shfl_xor(key, 1), k2 = __shfl_xor(key, 2), k3 = __shfl_xor(key, 3);
            const bool grp = (key == k1) & (key == k2) & (key == k3);
            long long iv[8], ivs[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                iv[q] = fx(v[q], fs);
                ivs[q] = iv[q] + __shfl_xor(iv[q], 1);
                ivs[q] += __shfl_xor(ivs[q], 2);
            }
            if (!act || (grp && (t & 3) != 0)) continue;
            const long long *val = grp ? ivs : iv;
            const int cx[4] = {ix, ix + 1, ix, ix + 1};
            const int cy[4] = {iy, iy, iy + 1, iy + 1};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (cy[k] < r0 || cy[k] > r1 || cx[k] > bx1) continue;
                int lk = (cy[k] - r0) * ww + (cx[k] - bx0);
                fx_add(&tile[2 * lk], val[2 * k]);
                fx_add(&tile[2 * lk + 1], val[2 * k + 1]);
            }
        }
        __syncthreads();
        for (int k = t; k < ncell; k += nt) {
            int ry = r0 + k / ww, rx = bx0 + k % ww;
            plane[ry * ng + rx] = make_double2((double)tile[2 * k] / fs, (double)tile[2 * k + 1] / fs);
        }
        __syncthreads();
    }
}

// Deterministic block sum of K doubles for any multiple of 64 threads:
// butterfly within each wavefront, then the wave totals in wave order.
template <int NT, int K>
__device__ __forceinline__ void block_sum_waves(double (&v)[K], double *lds) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) lds[wv * K + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        double acc = lds[k];
        for (int q = 1; q < NT / 64; q++) acc += lds[q * K + k];
        v[k] = acc;
    }
    __syncthreads();
}

// Block minimum of K ints, same structure; lds holds K * NT/64 ints.
template <int NT, int K>
__device__ __forceinline__ void block_min_waves(int (&v)[K], int *lds) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
#pragma unroll
    for (int k = 0; k < K; k++)
        for (int o = 32; o > 0; o >>= 1) v[k] = min(v[k], __shfl_xor(v[k], o));
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) lds[wv * K + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        int acc = lds[k];
        for (int q = 1; q < NT / 64; q++) acc = min(acc, lds[q * K + k]);
        v[k] = acc;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// The FP64 step's particle passes in two kernels, one workgroup per slice.
//
// front: first transverse half step (k_transverse), the slice moments
//   (k_slice_moments, same per-lane order and block tree, so the same bits),
//   sizes and centroid (k_slice_sizes), the largest radius from the
//   centroid (k_sc_rbound; a second read of the slice, from L2) and the
//   footprint box of the in-grid particles for the main kernel's tile.
// main: the short-range solve (k_sc_solve's histograms, operator and
//   recurrence, one mode solution per (m, l) kept in LDS), then one pass
//   per particle: ez from those modes, the RK4 push (k_push), R56 (k_r56),
//   the second transverse half step, the beam sums of DiagBeam (k_diag_beam,
//   same order) and the deposit (k_deposit) into an LDS tile over the
//   footprint padded by 2 cells; a slice whose particles leave the tile
//   takes the exact two-pass deposit over its box instead.
// Every value is computed by the same expression as in the separate
// kernels; only LDS-atomic summation orders of the histograms and the
// deposit differ, as between any two runs of those kernels.
// ---------------------------------------------------------------------------
constexpr int kFrontNT = 256;                          // the front's block tree is k_slice_moments'
constexpr int kFuseNT = 512;
constexpr int kFuseNW = kFuseNT / 64;
constexpr int kFuseModes = 512;                        // mode-solution cells (double2)
constexpr int kFuseDyn = kFuseModes + 3584;            // double2 cells of the dynamic region (64 KiB)
constexpr int kFuseTile = kFuseDyn - kFuseModes;        // deposit tile cells
static_assert(kFuseNW * kScMaxNg + 2 * kScMaxNg + (6 * kScMaxNg + 2) / 2 <= kFuseTile,
              "the short-range solver's scratch must fit in the tile region");

__global__ void __launch_bounds__(kFrontNT) k_fp64_front(const long long *__restrict__ off,
                                                        const double *__restrict__ g, double *__restrict__ x,
                                                        double *__restrict__ y, double *__restrict__ px,
                                                        double *__restrict__ py, const double *__restrict__ cur,
                                                        TransParams t1, int moments, int sc, double gridmax,
                                                        double dgrid, int ng, double *__restrict__ mom,
                                                        double *__restrict__ fsize, double *__restrict__ cen,
                                                        double *__restrict__ rbound, int4 *__restrict__ pbox,
                                                        DepositParams d, double *__restrict__ pmaxf) {
    __shared__ double lds[4 * kFrontNT];
    __shared__ int red[4][kFrontNT];
    const int is = blockIdx.x, t = threadIdx.x;
    const long long a = off[is], b = off[is + 1];
    double v[4] = {0, 0, 0, 0};
    int bx0 = 1 << 30, by0 = 1 << 30, bx1 = -1, by1 = -1;
    double pmax = 0;  // largest sqrt(faw2) / gamma of an in-grid particle (the deposit's scale)
    // the next particle's coordinates are loaded before this one's stores,
    // so two particles' loads are in flight per lane
    double nx = 0, ny = 0, npx = 0, npy = 0, ng_ = 0;
    if (a + t < b) {
        nx = x[a + t];
        ny = y[a + t];
        npx = px[a + t];
        npy = py[a + t];
        ng_ = g[a + t];
    }
    for (long long i = a + t; i < b; i += kFrontNT) {
        double xi = nx, yi = ny, pxi = npx, pyi = npy;
        const double gi = ng_;
        if (i + kFrontNT < b) {
            nx = x[i + kFrontNT];
            ny = y[i + kFrontNT];
            npx = px[i + kFrontNT];
            npy = py[i + kFrontNT];
            ng_ = g[i + kFrontNT];
        }
        transverse_one(t1, gi, xi, yi, pxi, pyi);
        x[i] = xi;
        y[i] = yi;
        px[i] = pxi;
        py[i] = pyi;
        v[0] += xi;
        v[1] += xi * xi;
        v[2] += yi;
        v[3] += yi * yi;
        if ((xi > -gridmax) && (xi < gridmax) && (yi > -gridmax) && (yi < gridmax)) {
            const int ix = min((int)floor((xi + gridmax) / dgrid), ng - 2);
            const int iy = min((int)floor((yi + gridmax) / dgrid), ng - 2);
            bx0 = min(bx0, ix);
            bx1 = max(bx1, ix + 1);
            by0 = min(by0, iy);
            by1 = max(by1, iy + 1);
            const double ddx = xi - d.ax, ddy = yi - d.ay;
            const double faw2 = (1 + d.kx * ddx * ddx + d.ky * ddy * ddy + 2 * (d.gradx * ddx + d.grady * ddy));
            pmax = fmax(pmax, sqrt(faw2) / fabs(gi));
        }
    }
    red[0][t] = bx0;
    red[1][t] = by0;
    red[2][t] = -bx1;
    red[3][t] = -by1;
    lds[t] = pmax;
    __syncthreads();
    for (int q = kFrontNT / 2; q > 0; q >>= 1) {
        if (t < q) {
            for (int k = 0; k < 4; k++) red[k][t] = min(red[k][t], red[k][t + q]);
            lds[t] = fmax(lds[t], lds[t + q]);
        }
        __syncthreads();
    }
    if (t == 0) {
        pbox[is] = make_int4(red[0][0], red[1][0], -red[2][0], -red[3][0]);
        pmaxf[is] = lds[0];
    }
    if (!moments) return;  // uniform
    block_sum<kFrontNT, 4>(v, lds);
    const long long cnt = b - a;
    double xc = 0, yc = 0;
    if (cnt > 0) {
        xc = v[0] / (double)cnt;
        yc = v[2] / (double)cnt;
    }
    if (t == 0) {
        for (int k = 0; k < 4; k++) mom[4 * is + k] = v[k];
        double sz;
        if (cur[is] <= 0 || cnt == 0) {
            sz = 1;
        } else {
            const double c = (double)cnt;
            const double x1 = v[0] / c, x2 = v[1] / c, y1 = v[2] / c, y2 = v[3] / c;
            sz = sqrt(fabs(x2 - x1 * x1)) * sqrt(fabs(y2 - y1 * y1));
        }
        if (sz <= 0) sz = 1;
        fsize[is] = sz;
        cen[2 * is] = xc;
        cen[2 * is + 1] = yc;
    }
    if (!sc) return;  // uniform
    double m = 0;
    for (long long i = a + t; i < b; i += kFrontNT) {
        const double tx = x[i] - xc, ty = y[i] - yc;
        const double r2 = tx * tx + ty * ty;
        if (r2 > m) m = r2;
    }
    lds[t] = m;
    __syncthreads();
    for (int q = kFrontNT / 2; q > 0; q >>= 1) {
        if (t < q) lds[t] = fmax(lds[t], lds[t + q]);
        __syncthreads();
    }
    if (t == 0) rbound[is] = sqrt(lds[0]);
}


__global__ void __launch_bounds__(kFuseNT) k_fp64_main(const long long *__restrict__ off, double *__restrict__ g,
                                                       double *__restrict__ th, double *__restrict__ x,
                                                       double *__restrict__ y, double *__restrict__ px,
                                                       double *__restrict__ py, const double2 *__restrict__ field,
                                                       const double *__restrict__ esc, const double *__restrict__ cur,
                                                       const double *__restrict__ cen, const double *__restrict__ dr_arr,
                                                       double *__restrict__ efield, const int4 *__restrict__ pbox,
                                                       const double *__restrict__ pmaxf, double2 *__restrict__ src,
                                                       int4 *__restrict__ box, double *__restrict__ beamsum,
                                                       FuseParams f) {
    extern __shared__ double2 fdyn[];
    __shared__ double wred[9 * kFuseNW];
    __shared__ int ired[5 * kFuseNW];
    double2 *celm = fdyn;                                      // [modes][ngr]
    double2 *work = fdyn + kFuseModes;                         // solver scratch, then the deposit tile
    const int is = blockIdx.x, t = threadIdx.x, w = t >> 6;
    const long long a = off[is], b = off[is + 1];
    const int npart = (int)(b - a);
    const PushParams &p = f.p;
    const DepositParams &d = f.d;
    const int ngr = f.ngr, ng = p.ngrid;
    const bool sc = f.use_sc && npart > 0;
    double xc = 0, yc = 0, dr = 1;
    if (t == 0 && f.use_sc) efield[is] = 0;
    // ---- short-range space charge (k_sc_solve): every mode's solution into celm
    if (sc) {
        long long(*hist)[kScMaxNg][2] = (long long(*)[kScMaxNg][2])work;   // [waves][ngr][2] fixed point
        double2 *gam = work + kFuseNW * kScMaxNg, *rhs = gam + kScMaxNg;
        double *vol = (double *)(rhs + kScMaxNg), *ldig = vol + kScMaxNg, *rlog = ldig + kScMaxNg + 1;
        double *cupp = rlog + kScMaxNg, *clow = cupp + kScMaxNg, *cmid = clow + kScMaxNg;
        dr = dr_arr[is];
        xc = cen[2 * is];
        yc = cen[2 * is + 1];
        const double pi = 2 * asin(1.0);
        for (int j = t; j <= ngr; j += kFuseNT) {
            if (j == ngr) {
                ldig[ngr] = 0;
            } else if (j == 0) {
                vol[0] = pi * dr * dr;
                rlog[0] = 0.5;
                ldig[0] = 0;
            } else {
                vol[j] = pi * dr * dr * (2 * j + 1);
                ldig[j] = 2 * pi * j;
                rlog[j] = log((double)(j + 1) / (double)j);
            }
        }
        const double coef = -f.gz2 / f.ks / f.ks;
        const double econst = kVacimp / kEev * cur[is] / f.ks / (double)npart;
        const double hs = fx_scale((double)npart);
        int mode = 0;
        for (int m = -f.nphi; m <= f.nphi; m++) {
            for (int l = 1; l <= f.nz; l++, mode++) {
                __syncthreads();
                for (int k = t; k < kFuseNW * kScMaxNg; k += kFuseNT)
                    hist[k / kScMaxNg][k % kScMaxNg][0] = hist[k / kScMaxNg][k % kScMaxNg][1] = 0;
                __syncthreads();
                for (long long i = a + t; i < b; i += kFuseNT) {
                    double tx = x[i] - xc, ty = y[i] - yc;
                    double radi = sqrt(tx * tx + ty * ty);
                    int bin = (int)floor(radi / dr);
                    bin = (bin >= 0 && bin < ngr) ? bin : (bin >= ngr ? ngr - 1 : 0);
                    double s, c;
                    sincos(l * th[i], &s, &c);
                    double2 v = make_double2(c, -s);
                    if (m != 0) v = cmul(cpow_int(make_double2(tx / radi, ty / radi), -m), v);
                    fx_add(&hist[w][bin][0], fx(v.x, hs));
                    fx_add(&hist[w][bin][1], fx(v.y, hs));
                }
                __syncthreads();
                for (int i = t; i < ngr; i += kFuseNT) {
                    double lm = -ldig[i] - ldig[i + 1] - 2. * pi * m * m * rlog[i];
                    if (i == ngr - 1) lm -= 2 * pi * ngr;
                    cmid[i] = 1 + coef * lm / l / l / vol[i];
                    if (i > 0) {
                        cupp[i] = coef * ldig[i] / l / l / vol[i - 1];
                        clow[i] = coef * ldig[i] / l / l / vol[i];
                    }
                    long long csx = hist[0][i][0], csy = hist[0][i][1];
                    for (int q = 1; q < kFuseNW; q++) {
                        csx += hist[q][i][0];
                        csy += hist[q][i][1];
                    }
                    const double2 cs = make_double2((double)csx / hs, (double)csy / hs);
                    rhs[i] = cmul(cs, make_double2(0, econst / l / vol[i]));
                }
                __syncthreads();
                if (t == 0) {
                    double2 *ce = celm + mode * ngr;
                    double bet = cmid[0], inv = 1.0 / bet;
                    ce[0] = cscale(rhs[0], inv);
                    for (int i = 1; i < ngr; i++) {
                        const double gg = cupp[i] * inv;
                        gam[i].x = gg;
                        bet = cmid[i] - clow[i] * gg;
                        inv = 1.0 / bet;
                        ce[i] = cscale(csub(rhs[i], cscale(ce[i - 1], clow[i])), inv);
                    }
                    for (int i = ngr - 2; i > -1; i--) ce[i] = csub(ce[i], cscale(ce[i + 1], gam[i + 1].x));
                    if (l == 1) efield[is] = hypot(ce[0].x, ce[0].y);
                }
            }
        }
    }
    // ---- deposit tile over the footprint (front kernel) padded by 2 cells
    const bool dep = d.enabled != 0;
    const int4 pb = pbox[is];
    int tx0 = 0, ty0 = 0, tw = 0, tht = 0;
    bool fused = false;
    if (dep && pb.z >= pb.x) {
        tx0 = max(0, pb.x - 2);
        ty0 = max(0, pb.y - 2);
        tw = min(ng - 1, pb.z + 2) - tx0 + 1;
        tht = min(ng - 1, pb.w + 2) - ty0 + 1;
        fused = tw * tht <= kFuseTile;
    }
    double scl = 0;
    if (dep) {
        const double scl0 = d.fcoup * kVacimp * cur[is] * d.xks * d.delz;
        scl = scl0 / (4 * kEev * (double)(b - a) * d.dgrid * d.dgrid);
    }
    // fixed-point tile: a cell sums at most sum_i |part_i| <= n max|part|; the
    // front's max|part| (before the push) is doubled as a margin for the
    // step's energy change, and the pass checks that its own sum of |part|
    // (v[8]) stays within that bound, else it deposits again in two passes
    const double bound1 = 2.0 * (double)npart * pmaxf[is] * fabs(scl);
    const double fs1 = fx_scale(bound1);
    // the tile as two planes (real parts, then imaginary parts): a lane's
    // 8-byte atomic then spans two LDS banks, not four
    long long *itile = (long long *)work;  // [2][cells]
    __syncthreads();  // solver scratch dead: the tile takes its place
    if (fused)
        for (int k = t; k < 2 * tw * tht; k += kFuseNT) itile[k] = 0;
    long long *tre = itile, *tim = itile + tw * tht;
    __syncthreads();
    // ---- one pass per particle
    const double eloss = -esc[is] / kEev;
    const int islice = (is + p.first) % p.nslice;
    const double2 *fsl = field + (long long)islice * ng * ng;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // beam sums; [8]: sum of |part| (the tile's bound check)
    int bx0 = 1 << 30, by0 = 1 << 30, bx1 = -1, by1 = -1, ovf = 0;
    for (long long i = a + t; i < b; i += kFuseNT) {
        double gamma = g[i];
        double theta = th[i];
        double xi = x[i], yi = y[i], pxi = px[i], pyi = py[i];
        // short-range field of the particle (k_sc_solve's gather, mode order)
        double ez = 0;
        if (sc) {
            const double tx = xi - xc, ty = yi - yc;
            const double radi = sqrt(tx * tx + ty * ty);
            int bin = (int)floor(radi / dr);
            bin = (bin >= 0 && bin < ngr) ? bin : (bin >= ngr ? ngr - 1 : 0);
            int mode = 0;
            for (int m = -f.nphi; m <= f.nphi; m++) {
                for (int l = 1; l <= f.nz; l++, mode++) {
                    double s, c;
                    sincos(l * theta, &s, &c);
                    double2 ct = make_double2(c, s);
                    if (m != 0) ct = cmul(cpow_int(make_double2(tx / radi, ty / radi), m), ct);
                    ez += 2 * cmul(ct, celm[mode * ngr + bin]).x;
                }
            }
        }
        // push (k_push)
        theta += p.autophase;
        const double dx = xi - p.ax, dy = yi - p.ay;
        const double awloc = (1 + 0.5 * (p.kx * dx * dx + p.ky * dy * dy) + p.gradx * dx + p.grady * dy);
        Ode<0> o;
        o.xks = p.xks;
        o.xku = p.xku;
        o.btpar = 1 + pxi * pxi + pyi * pyi + p.aw * p.aw * awloc * awloc;
        o.ez = (p.use_sc ? ez : 0.0) + eloss;
        if ((xi > -p.gridmax) && (xi < p.gridmax) && (yi > -p.gridmax) && (yi < p.gridmax)) {
            double wx = (xi + p.gridmax) / p.dgrid;
            double wy = (yi + p.gridmax) / p.dgrid;
            int ix = (int)floor(wx), iy = (int)floor(wy);
            wx = 1 + floor(wx) - wx;
            wy = 1 + floor(wy) - wy;
            ix = min(ix, p.ngrid - 2);
            iy = min(iy, p.ngrid - 2);
            const int idx = ix + iy * p.ngrid;
            const double2 f0 = fsl[idx], f1 = fsl[idx + 1], f2 = fsl[idx + p.ngrid], f3 = fsl[idx + p.ngrid + 1];
            double2 cpart = cscale(cscale(f0, wx), wy);
            cpart = cadd(cpart, cscale(cscale(f1, 1 - wx), wy));
            cpart = cadd(cpart, cscale(cscale(f2, wx), 1 - wy));
            cpart = cadd(cpart, cscale(cscale(f3, 1 - wx), 1 - wy));
            const double r = p.rtmp * awloc;
            o.rpart = make_double2(r * cpart.x, r * -cpart.y);
        } else {
            o.rpart = make_double2(0, 0);
        }
        const double delz = p.delz;
        double k3gg, k3pp;
        o.k2gg = 0;
        o.k2pp = 0;
        o.eval(gamma, theta);
        double stpz = 0.5 * delz;
        gamma += stpz * o.k2gg;
        theta += stpz * o.k2pp;
        k3gg = o.k2gg;
        k3pp = o.k2pp;
        o.k2gg = 0;
        o.k2pp = 0;
        o.eval(gamma, theta);
        gamma += stpz * (o.k2gg - k3gg);
        theta += stpz * (o.k2pp - k3pp);
        k3gg /= 6;
        k3pp /= 6;
        o.k2gg *= -0.5;
        o.k2pp *= -0.5;
        o.eval(gamma, theta);
        stpz = delz;
        gamma += stpz * o.k2gg;
        theta += stpz * o.k2pp;
        k3gg -= o.k2gg;
        k3pp -= o.k2pp;
        o.k2gg *= 2;
        o.k2pp *= 2;
        o.eval(gamma, theta);
        gamma += stpz * (k3gg + o.k2gg / 6.0);
        theta += stpz * (k3pp + o.k2pp / 6.0);
        if (f.chicane) theta += f.r56 * (gamma - f.gamma0);   // k_r56
        transverse_one(f.t2, gamma, xi, yi, pxi, pyi);
        g[i] = gamma;
        th[i] = theta;
        x[i] = xi;
        y[i] = yi;
        px[i] = pxi;
        py[i] = pyi;
        // beam sums (k_diag_beam)
        v[0] += gamma;
        v[1] += gamma * gamma;
        v[2] += xi;
        v[3] += xi * xi;
        v[4] += yi;
        v[5] += yi * yi;
        double s1, c1;
        sincos((0 + 1) * theta, &s1, &c1);
        v[6] += c1;
        v[7] += s1;
        // deposit (k_deposit)
        if (dep && (xi > -d.gridmax) && (xi < d.gridmax) && (yi > -d.gridmax) && (yi < d.gridmax)) {
            double wx = (xi + d.gridmax) / d.dgrid;
            double wy = (yi + d.gridmax) / d.dgrid;
            const int ix = min((int)floor(wx), ng - 2), iy = min((int)floor(wy), ng - 2);
            bx0 = min(bx0, ix);
            bx1 = max(bx1, ix + 1);
            by0 = min(by0, iy);
            by1 = max(by1, iy + 1);
            const double ddx = xi - d.ax, ddy = yi - d.ay;
            const double faw2 = (1 + d.kx * ddx * ddx + d.ky * ddy * ddy + 2 * (d.gradx * ddx + d.grady * ddy));
            const double part = sqrt(faw2) * scl / gamma;
            v[8] += fabs(part);
            if (fused) {
                if (ix >= tx0 && ix + 1 < tx0 + tw && iy >= ty0 && iy + 1 < ty0 + tht) {
                    wx = 1 + floor(wx) - wx;
                    wy = 1 + floor(wy) - wy;
                    double s, c;
                    sincos(d.harm * theta, &s, &c);
                    const double2 cp = make_double2(s * part, c * part);
                    const int lk = (iy - ty0) * tw + (ix - tx0);
                    const double w00 = wx * wy, w10 = (1 - wx) * wy, w01 = wx * (1 - wy), w11 = (1 - wx) * (1 - wy);
                    fx_add(&tre[lk], fx(w00 * cp.x, fs1));
                    fx_add(&tim[lk], fx(w00 * cp.y, fs1));
                    fx_add(&tre[lk + 1], fx(w10 * cp.x, fs1));
                    fx_add(&tim[lk + 1], fx(w10 * cp.y, fs1));
                    fx_add(&tre[lk + tw], fx(w01 * cp.x, fs1));
                    fx_add(&tim[lk + tw], fx(w01 * cp.y, fs1));
                    fx_add(&tre[lk + tw + 1], fx(w11 * cp.x, fs1));
                    fx_add(&tim[lk + tw + 1], fx(w11 * cp.y, fs1));
                } else {
                    ovf = 1;
                }
            }
        }
    }
    block_sum_waves<kFuseNT, 9>(v, wred);
    if (t == 0)
        for (int k = 0; k < 8; k++) beamsum[8 * is + k] = v[k];
    if (!(v[8] <= bound1)) ovf = 1;  // a particle's coupling grew past the margin: exact two-pass deposit
    // ---- footprint box and overflow flag
    int bv[5] = {bx0, by0, -bx1, -by1, -ovf};
    block_min_waves<kFuseNT, 5>(bv, ired);
    bx0 = bv[0];
    by0 = bv[1];
    bx1 = -bv[2];
    by1 = -bv[3];
    ovf = -bv[4];
    if (!dep || bx1 < 0) {  // no source this step: an empty box
        if (t == 0) box[is] = make_int4(1, 1, 0, 0);
        return;
    }
    double2 *plane = src + (long long)is * ng * ng;
    if (fused && !ovf) {
        for (int k = t; k < tw * tht; k += kFuseNT)
            plane[(ty0 + k / tw) * ng + tx0 + k % tw] =
                make_double2((double)tre[k] / fs1, (double)tim[k] / fs1);
        if (t == 0) box[is] = make_int4(tx0, ty0, tx0 + tw - 1, ty0 + tht - 1);
        return;
    }
    const double fs2 = fx_scale(v[8]);  // the pass's own sum of |part|: exact bound
    // ---- exact two-pass deposit over the footprint box (rows in bands of the tile)
    bx1 = min(bx1, ng - 1);
    by1 = min(by1, ng - 1);
    if (t == 0) box[is] = make_int4(bx0, by0, bx1, by1);
    const int ww = bx1 - bx0 + 1;
    const int rows_per_band = max(1, kFuseTile / ww);
    for (int r0 = by0; r0 <= by1; r0 += rows_per_band) {
        const int r1 = min(by1, r0 + rows_per_band - 1);
        const int ncell = (r1 - r0 + 1) * ww;
        for (int k = t; k < 2 * ncell; k += kFuseNT) itile[k] = 0;
        long long *pre = itile, *pim = itile + ncell;  // two planes, as the fused tile
        __syncthreads();
        for (long long i = a + t; i < b; i += kFuseNT) {
            const double xi = x[i], yi = y[i];
            if (!((xi > -d.gridmax) && (xi < d.gridmax) && (yi > -d.gridmax) && (yi < d.gridmax))) continue;
            double wx = (xi + d.gridmax) / d.dgrid;
            double wy = (yi + d.gridmax) / d.dgrid;
            const int ix = min((int)floor(wx), ng - 2), iy = min((int)floor(wy), ng - 2);
            if (iy + 1 < r0 || iy > r1) continue;
            wx = 1 + floor(wx) - wx;
            wy = 1 + floor(wy) - wy;
            const double ddx = xi - d.ax, ddy = yi - d.ay;
            const double faw2 = (1 + d.kx * ddx * ddx + d.ky * ddy * ddy + 2 * (d.gradx * ddx + d.grady * ddy));
            const double part = sqrt(faw2) * scl / g[i];
            double s, c;
            sincos(d.harm * th[i], &s, &c);
            const double2 cp = make_double2(s * part, c * part);
            const double wgt[4] = {wx * wy, (1 - wx) * wy, wx * (1 - wy), (1 - wx) * (1 - wy)};
            const int cx[4] = {ix, ix + 1, ix, ix + 1};
            const int cy[4] = {iy, iy, iy + 1, iy + 1};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (cy[k] < r0 || cy[k] > r1 || cx[k] < bx0 || cx[k] > bx1) continue;
                const int lk = (cy[k] - r0) * ww + (cx[k] - bx0);
                fx_add(&pre[lk], fx(wgt[k] * cp.x, fs2));
                fx_add(&pim[lk], fx(wgt[k] * cp.y, fs2));
            }
        }
        __syncthreads();
        for (int k = t; k < ncell; k += kFuseNT) {
            const int ry = r0 + k / ww, rx = bx0 + k % ww;
            plane[ry * ng + rx] = make_double2((double)pre[k] / fs2, (double)pim[k] / fs2);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// K5/K6: ADI half steps, FieldSolverADI::ADI (:60-91) with tridagx
// (:94-104) and tridagy (:106-122).  One lane per line; the x sweep writes
// the intermediate field into `work`, the y sweep writes the result back
// into the field slice.  Beam slice ii drives field slice (ii+first)%ns.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 src_at(const double2 *plane, int4 bx, int ix, int iy, int ng) {
    if (ix < bx.x || ix > bx.z || iy < bx.y || iy > bx.w) return make_double2(0, 0);
    return plane[iy * ng + ix];
}

__global__ void __launch_bounds__(64) k_adi_x(double2 *__restrict__ field,
                                              double2 *__restrict__ work,
                                              const double2 *__restrict__ src,
                                              const int4 *__restrict__ box, AdiFactors f,
                                              int first, int ns) {
    const int ng = f.ngrid;
    const int iy = blockIdx.x * 64 + threadIdx.x;
    const int ii = blockIdx.y;
    if (iy >= ng) return;
    const long long fs = (long long)((ii + first) % ns) * ng * ng;
    const double2 *u = field + fs;
    double2 *w = work + fs;
    const double2 *S = src + (long long)ii * ng * ng;
    const int4 bx = box[ii];
    const double2 cst = make_double2(0.0, f.cstep);
    double2 yprev = make_double2(0, 0);
    for (int k = 0; k < ng; k++) {
        const int idx = iy * ng + k;
        double2 uc = u[idx];
        double2 st;
        if (iy == 0) {
            st = csub(u[idx + ng], cscale(uc, 2.0));
        } else if (iy == ng - 1) {
            st = csub(u[idx - ng], cscale(uc, 2.0));
        } else {
            st = cadd(csub(u[idx + ng], cscale(uc, 2.0)), u[idx - ng]);
        }
        double2 r = cadd(cadd(src_at(S, bx, k, iy, ng), uc), cmul(cst, st));
        double2 yk = (k == 0) ? cmul(r, f.cbet[0]) : cmul(csub(r, cmul(f.c[k], yprev)), f.cbet[k]);
        w[idx] = yk;
        yprev = yk;
    }
    for (int k = ng - 2; k >= 0; k--) {
        const int idx = iy * ng + k;
        double2 yk = csub(w[idx], cmul(f.cwet[k + 1], yprev));
        w[idx] = yk;
        yprev = yk;
    }
}

__global__ void __launch_bounds__(64) k_adi_y(double2 *__restrict__ field,
                                              const double2 *__restrict__ work,
                                              const double2 *__restrict__ src,
                                              const int4 *__restrict__ box, AdiFactors f,
                                              int first, int ns) {
    const int ng = f.ngrid;
    const int ix = blockIdx.x * 64 + threadIdx.x;
    const int ii = blockIdx.y;
    if (ix >= ng) return;
    const long long fs = (long long)((ii + first) % ns) * ng * ng;
    double2 *u = field + fs;
    const double2 *w = work + fs;
    const double2 *S = src + (long long)ii * ng * ng;
    const int4 bx = box[ii];
    const double2 cst = make_double2(0.0, f.cstep);
    double2 yprev = make_double2(0, 0);
    for (int k = 0; k < ng; k++) {
        const int idx = k * ng + ix;
        double2 wc = w[idx];
        double2 st;
        if (ix == 0) {
            st = csub(w[idx + 1], cscale(wc, 2.0));
        } else if (ix == ng - 1) {
            st = csub(w[idx - 1], cscale(wc, 2.0));
        } else {
            st = cadd(csub(w[idx + 1], cscale(wc, 2.0)), w[idx - 1]);
        }
        double2 r = cadd(cadd(src_at(S, bx, ix, k, ng), wc), cmul(cst, st));
        double2 yk = (k == 0) ? cmul(r, f.cbet[0]) : cmul(csub(r, cmul(f.c[k], yprev)), f.cbet[k]);
        u[idx] = yk;
        yprev = yk;
    }
    for (int k = ng - 2; k >= 0; k--) {
        const int idx = k * ng + ix;
        double2 yk = csub(u[idx], cmul(f.cwet[k + 1], yprev));
        u[idx] = yk;
        yprev = yk;
    }
}

// ---------------------------------------------------------------------------
// K5+K6 fused: both ADI half steps of one slice per workgroup.
//
// FieldSolverADI::ADI (:60-91) solves (1 - i a d_x^2) u' = S + (1 + i a d_y^2) u
// row by row, then (1 - i a d_y^2) u'' = S + (1 + i a d_x^2) u' column by
// column, with Thomas' algorithm (tridagx :94-104, tridagy :106-122).
// Each lane owns one line and streams along it in tiles of T cells:
//  * forward elimination y_k = (r_k - c_k y_{k-1}) cbet_k exactly as the
//    reference;
//  * back substitution u_k = y_k - cwet_{k+1} u_{k+1} run over a two-tile
//    register window, started from y at the window's far end.  The start
//    error reaches tile cells only through a product of >= T+1 factors
//    |cwet| (the host checks max|cwet|^(T+1) < 2^-66 before choosing this
//    kernel), so the result equals Thomas' to rounding.  At the line's end
//    the start value is exact.
// Row tiles of the x sweep are staged through LDS so every global access is
// a contiguous run of T cells; the intermediate u' goes to `work` (it stays
// in L2/MALL between the sweeps of a slice).  The y sweep also accumulates
// the field diagnostics of DiagField::getValues (:763-861) on the final
// values: diag[8*fs + {power, x1, x2, y1, y2, ff_re, ff_im, on-axis |E|^2}].
// ---------------------------------------------------------------------------
// Barrier that orders LDS traffic only: outstanding global stores are not
// waited for (the sweeps' global results are published once, at the end of
// the x sweep, with a full __syncthreads()).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}


template <int T, int G, bool PREF>
__global__ void __launch_bounds__(G <= 192 ? 192 : 256)
    k_adi_fused(double2 *__restrict__ field, double2 *__restrict__ work,
                const double2 *__restrict__ src, const int4 *__restrict__ box, AdiFactors f,
                int first, int ns, double *__restrict__ diag) {
    constexpr int NT = G <= 192 ? 192 : 256;
    // prefetch registers per lane (ng <= G); without PREF tiles are loaded
    // straight into LDS when they are needed
    constexpr int PF = PREF ? (G * T + NT - 1) / NT : 1;
    extern __shared__ double2 lds[];
    const int ng = f.ngrid;
    const int t = threadIdx.x;
    const int ii = blockIdx.x;
    const long long fs = (long long)((ii + first) % ns);
    double2 *u = field + fs * ng * ng;
    double2 *w = work + fs * ng * ng;
    const double2 *S = src + (long long)ii * ng * ng;
    const int4 bx = box[ii];
    const double2 cst = make_double2(0.0, f.cstep);
    const int ntile = (ng + T - 1) / T;
    const int tcells = ng * T;
    // x-sweep tiles are [row][T] with a row pitch of T+1 cells: lane iy
    // reads row iy, and a pitch of 4 (mod 32) dwords keeps each group of 8
    // lanes of a 16-byte LDS read on disjoint banks (pitch T would put the
    // whole wavefront on the same four banks).
    constexpr int TP = T + 1;
    const int tstride = ng * TP;  // >= tcells, also holds the y-sweep tiles
    // tile buffers addressed by index so every access stays an LDS access
    auto buf = [&](int q) { return lds + (q & 1) * tstride; };
    // Thomas factors in LDS: fc[k], fb[k], fw[k]
    double2 *fc = lds + 2 * tstride, *fb = fc + ng, *fw = fb + ng;
    for (int k = t; k < ng; k += NT) {
        fc[k] = f.c[k];
        fb[k] = f.cbet[k];
        fw[k] = f.cwet[k];
    }
    double2 ya[T], yb[T], pf[PF], sv[T];
    double2 yprev = make_double2(0, 0);

    // ---------------- x sweep: lane = row iy, stream along ix -------------
    // tile j holds columns [jT, jT+T) of every row: buf[.][row*T + col]
    auto fetch_x = [&](int j) {
#pragma unroll
        for (int q = 0; q < PF; q++) {
            const int e = t + q * NT;
            const int row = e / T, col = j * T + e % T;
            pf[q] = (e < tcells && col < ng) ? u[row * ng + col] : make_double2(0, 0);
        }
    };
    auto stash_x = [&](double2 *dst) {
#pragma unroll
        for (int q = 0; q < PF; q++) {
            const int e = t + q * NT;
            if (e < tcells) dst[(e / T) * TP + e % T] = pf[q];
        }
    };
    auto direct_x = [&](int j, double2 *dst) {
        for (int e = t; e < tcells; e += NT) {
            const int row = e / T, col = j * T + e % T;
            dst[row * TP + e % T] = col < ng ? u[row * ng + col] : make_double2(0, 0);
        }
    };
    if (PREF) {
        fetch_x(0);
        stash_x(buf(0));
    } else {
        direct_x(0, buf(0));
    }
    __syncthreads();
    const int iy = t;
    const bool srow = iy >= bx.y && iy <= bx.w;  // this row crosses the source box
    for (int j = 0; j <= ntile; j++) {
        double2 *cur = buf(j);
        // source cells of this tile first, then the next tile's prefetch,
        // so waiting for the former never waits for the latter
#pragma unroll
        for (int kk = 0; kk < T; kk++) {
            const int k = j * T + kk;
            sv[kk] = (j < ntile && iy < ng && srow && k >= bx.x && k <= bx.z && k < ng)
                         ? S[iy * ng + k] : make_double2(0, 0);
        }
        if (PREF && j + 1 < ntile) fetch_x(j + 1);
        if (iy < ng && j < ntile) {
#pragma unroll
            for (int kk = 0; kk < T; kk++) {
                const int k = j * T + kk;
                if (k < ng) {
                    const double2 uc = cur[iy * TP + kk];
                    double2 st;
                    if (iy == 0) st = csub(cur[(iy + 1) * TP + kk], cscale(uc, 2.0));
                    else if (iy == ng - 1) st = csub(cur[(iy - 1) * TP + kk], cscale(uc, 2.0));
                    else st = cadd(csub(cur[(iy + 1) * TP + kk], cscale(uc, 2.0)), cur[(iy - 1) * TP + kk]);
                    const double2 r = cadd(cadd(sv[kk], uc), cmul(cst, st));
                    yb[kk] = (k == 0) ? cmul(r, fb[0]) : cmul(csub(r, cmul(fc[k], yprev)), fb[k]);
                    yprev = yb[kk];
                }
            }
        }
        lds_barrier();  // every lane is done reading `cur`: reuse it for output
        if (iy < ng && j > 0) {
            const int kend = (j < ntile) ? min(j * T + T - 1, ng - 1) : (j * T - 1);
            double2 v = make_double2(0, 0);
            bool started = false;
            if (j < ntile) {
#pragma unroll
                for (int kk = T - 1; kk >= 0; kk--) {
                    const int k = j * T + kk;
                    if (k <= kend) {
                        v = started ? csub(yb[kk], cmul(fw[k + 1], v)) : yb[kk];
                        started = true;
                    }
                }
            }
#pragma unroll
            for (int kk = T - 1; kk >= 0; kk--) {
                const int k = (j - 1) * T + kk;
                if (k < ng) {
                    v = started ? csub(ya[kk], cmul(fw[k + 1], v)) : ya[kk];
                    started = true;
                    cur[iy * TP + kk] = v;
                }
            }
        }
        lds_barrier();
        if (j > 0) {
            const int k0 = (j - 1) * T;
            for (int e = t; e < tcells; e += NT) {
                const int row = e / T, col = k0 + e % T;
                if (col < ng) w[row * ng + col] = cur[row * TP + e % T];
            }
        }
        if (j + 1 < ntile) {
            if (PREF) stash_x(buf(j + 1));
            else direct_x(j + 1, buf(j + 1));
        }
#pragma unroll
        for (int kk = 0; kk < T; kk++) ya[kk] = yb[kk];
        lds_barrier();
    }
    __syncthreads();  // publish the intermediate field of the whole slice

    // ---------------- y sweep: lane = column ix, stream along iy ----------
    // tile j holds rows [jT, jT+T) of every column: buf[.][row*ng + col]
    auto fetch_y = [&](int j) {
#pragma unroll
        for (int q = 0; q < PF; q++) {
            const int e = t + q * NT;
            const int row = j * T + e / ng, col = e % ng;
            pf[q] = (e < tcells && row < ng) ? w[row * ng + col] : make_double2(0, 0);
        }
    };
    auto stash = [&](double2 *dst) {
#pragma unroll
        for (int q = 0; q < PF; q++) {
            const int e = t + q * NT;
            if (e < tcells) dst[e] = pf[q];
        }
    };
    auto direct_y = [&](int j, double2 *dst) {
        for (int e = t; e < tcells; e += NT) {
            const int row = j * T + e / ng, col = e % ng;
            dst[e] = row < ng ? w[row * ng + col] : make_double2(0, 0);
        }
    };
    const int ix = t;
    const bool scol = ix >= bx.x && ix <= bx.z;
    const double shift = -0.5 * (double)(ng - 1);
    double dg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int icen = (ng / 2) * ng + (ng / 2);
    yprev = make_double2(0, 0);
    if (PREF) {
        fetch_y(0);
        stash(buf(0));
    } else {
        direct_y(0, buf(0));
    }
    lds_barrier();
    for (int j = 0; j <= ntile; j++) {
        const double2 *cur = buf(j);
#pragma unroll
        for (int kk = 0; kk < T; kk++) {
            const int k = j * T + kk;
            sv[kk] = (j < ntile && ix < ng && scol && k >= bx.y && k <= bx.w && k < ng)
                         ? S[k * ng + ix] : make_double2(0, 0);
        }
        if (PREF && j + 1 < ntile) fetch_y(j + 1);
        if (ix < ng && j < ntile) {
#pragma unroll
            for (int kk = 0; kk < T; kk++) {
                const int k = j * T + kk;
                if (k < ng) {
                    const double2 wc = cur[kk * ng + ix];
                    double2 st;
                    if (ix == 0) st = csub(cur[kk * ng + ix + 1], cscale(wc, 2.0));
                    else if (ix == ng - 1) st = csub(cur[kk * ng + ix - 1], cscale(wc, 2.0));
                    else st = cadd(csub(cur[kk * ng + ix + 1], cscale(wc, 2.0)), cur[kk * ng + ix - 1]);
                    const double2 r = cadd(cadd(sv[kk], wc), cmul(cst, st));
                    yb[kk] = (k == 0) ? cmul(r, fb[0]) : cmul(csub(r, cmul(fc[k], yprev)), fb[k]);
                    yprev = yb[kk];
                }
            }
        }
        if (ix < ng && j > 0) {
            const int kend = (j < ntile) ? min(j * T + T - 1, ng - 1) : (j * T - 1);
            double2 v = make_double2(0, 0);
            bool started = false;
            if (j < ntile) {
#pragma unroll
                for (int kk = T - 1; kk >= 0; kk--) {
                    const int k = j * T + kk;
                    if (k <= kend) {
                        v = started ? csub(yb[kk], cmul(fw[k + 1], v)) : yb[kk];
                        started = true;
                    }
                }
            }
#pragma unroll
            for (int kk = T - 1; kk >= 0; kk--) {
                const int k = (j - 1) * T + kk;
                if (k < ng) {
                    v = started ? csub(ya[kk], cmul(fw[k + 1], v)) : ya[kk];
                    started = true;
                    u[k * ng + ix] = v;
                    const double wei = v.x * v.x + v.y * v.y;
                    const double dx = (double)ix + shift, dy = (double)k + shift;
                    dg[0] += wei;
                    dg[1] += dx * wei;
                    dg[2] += dx * dx * wei;
                    dg[3] += dy * wei;
                    dg[4] += dy * dy * wei;
                    dg[5] += v.x;
                    dg[6] += v.y;
                    if (k * ng + ix == icen) dg[7] = wei;
                }
            }
        }
        if (j + 1 < ntile) {
            if (PREF) stash(buf(j + 1));
            else direct_y(j + 1, buf(j + 1));
        }
#pragma unroll
        for (int kk = 0; kk < T; kk++) ya[kk] = yb[kk];
        lds_barrier();
    }
    // block reduction of the diagnostics (fixed order); the on-axis value
    // is carried by exactly one lane, others add 0
    double *red = (double *)lds;
    block_sum_waves<NT, 8>(dg, red);
    if (t == 0)
        for (int k = 0; k < 8; k++) diag[8 * fs + k] = dg[k];
}

// ---------------------------------------------------------------------------
// K5+K6 as a banded operator: both ADI half steps for a 32x32 output tile
// per workgroup.
//
// The x-implicit half step is u' = A^-1 (S + (1 + i a d_y^2) u) with A the
// constant tridiagonal matrix (1 - i a d^2) of FieldSolverADI::init
// (:125-182); the y half step is u'' = A^-1 (S + (1 + i a d_x^2) u').  The
// entries of A^-1 fall off like |a/(1+2a)|^|m| away from the diagonal, so
// A^-1 is, to rounding, the band |m| <= M whose coefficients band[k][m+M]
// = (A^-1)[k][k+m] the host builds from the same Thomas factors the
// reference uses and accepts only when every dropped entry is below 2^-60
// of the diagonal (1/128 ulp).  A tile then needs u on a halo of M+1 cells, and the
// update is a separable stencil: no sequential sweep, HBM traffic of one
// read and one write of the field (the result goes to the second field
// buffer; the host swaps them).  A is symmetric, so the rows of its inverse
// that equal the central (Toeplitz) row are symmetric too: interior tiles
// apply them as b_M E_k + sum_j b_{M+j} (E_{k-j} + E_{k+j}), M + 1 complex
// multiplies per output instead of 2M + 1.  Complex multiply-adds use fused FMA here:
// the rounding differs from Thomas' only at the last bits (the banded
// product is a different, equally accurate, evaluation of A^-1 r).
// Per-tile field-diagnostic sums go to tilediag[slice][tile][8] and are
// combined in tile order by k_adi_tile_diag.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 acc) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
    return acc;
}

template <int M>
struct BandGeom {
    static constexpr int TS = 32;                   // output tile edge
    static constexpr int W = 2 * M + 1;             // band width
    static constexpr int RE = TS + 2 * M;           // rows/cols of the E region
    // LDS row pitches of 1 (mod 8) cells (4 mod 32 dwords): the 8 lanes of a
    // 128-byte LDS access then fall on 8 distinct bank quads when they
    // walk down a column
    static constexpr int PE = ((RE + 6) / 8) * 8 + 1;
    static constexpr int PG = TS + 1;
    static constexpr int CX = 8;                    // outputs per lane in the x stage
    static constexpr int SR = (RE + (256 / RE) - 1) / (256 / RE);  // rows per lane in the E stage
    // E, then (after the x stage has read it) the y stage's right-hand side
    // F over the same cells with the y band rows of edge tiles behind it;
    // the x band rows of edge tiles in a region of their own
    static constexpr size_t lds_cells = (size_t)RE * PE + (size_t)TS * W;
    static_assert(PG <= PE, "F must fit in E's rows");
    static_assert(RE * (TS / CX) <= 256, "the x stage maps one lane per (row, CX columns): M <= 16");
    static_assert((size_t)RE * PG + (size_t)TS * W <= (size_t)RE * PE, "F and the y band must fit in E");
};

template <int M>
__global__ void __launch_bounds__(256) k_adi_band(const double2 *__restrict__ fin,
                                                  double2 *__restrict__ fout,
                                                  const double2 *__restrict__ src,
                                                  const int4 *__restrict__ box,
                                                  const double2 *__restrict__ band, double a,
                                                  int ng, int first, int ns, int ntx, int klo,
                                                  int khi, double *__restrict__ tilediag) {
    using Gm = BandGeom<M>;
    constexpr int TS = Gm::TS, W = Gm::W, RE = Gm::RE, PE = Gm::PE, PG = Gm::PG, CX = Gm::CX;
    constexpr int NT = 256;
    extern __shared__ double2 sm[];
    double2 *E = sm;                  // [RE][PE] E = S + (1 + i a d_y^2) u
    double2 *G = sm;                  // [RE][PG] F = S + 2 u' - E, written over E
    double2 *BY = sm + RE * PG;       // [TS][W] band rows of y (edge tiles), over E
    double2 *BX = sm + RE * PE;       // [TS][W] band rows of x (edge tiles)
    const int t = threadIdx.x;
    // XCD-aware order: workgroups are handed to the 8 XCDs round-robin, so
    // block b runs on XCD b % 8.  Blocks of a group of 8 slices are ordered
    // tile-major (b = group*8*ntiles + tile*8 + slice%8): all tiles of one
    // slice land on the same XCD, close in time, and share their halo rows
    // through that XCD's L2.
    const int ntiles = ntx * ntx;
    const int b = blockIdx.x;
    const int ii = (b / (8 * ntiles)) * 8 + (b & 7);   // beam slice
    const int tile = (b % (8 * ntiles)) >> 3;
    if (ii >= ns) return;  // uniform per block; padding of the last group
    const int tx0 = (tile % ntx) * TS, ty0 = (tile / ntx) * TS;
    const long long fs = (long long)((ii + first) % ns);
    const double2 *u = fin + fs * ng * ng;
    double2 *uo = fout + fs * ng * ng;
    const double2 *S = src + (long long)ii * ng * ng;
    const int4 bx = box[ii];
    auto sat = [&](int x, int y) {
        return (x >= bx.x && x <= bx.z && y >= bx.y && y <= bx.w) ? S[y * ng + x]
                                                                  : make_double2(0, 0);
    };
    // rows [klo, khi] of A^-1 share the band of row (klo+khi)/2 to 2^-60
    const bool xin = tx0 >= klo && tx0 + TS - 1 <= khi;
    const bool yin = ty0 >= klo && ty0 + TS - 1 <= khi;
    const double2 *bc = band + (long long)((klo + khi) / 2) * W;
    if (!xin)
        for (int e = t; e < TS * W; e += NT) {
            const int x = tx0 + e / W;
            BX[e] = x < ng ? band[(long long)x * W + e % W] : make_double2(0, 0);
        }
    // explicit y stencil of the x half step (FieldSolverADI.cpp:64-72),
    // zero outside the grid.  Lane -> one column of the E region and a strip
    // of SR rows: the SR+2 field loads (and the source loads) are all issued
    // before any is used, so a tile costs one memory round trip.
    {
        constexpr int NSTRIP = (RE + Gm::SR - 1) / Gm::SR;
        constexpr int SR = Gm::SR;
        if (t < RE * NSTRIP) {
            const int c = t % RE, r0 = (t / RE) * SR;
            const int x = tx0 - M + c;
            const bool xok = x >= 0 && x < ng;
            double2 uv[SR + 2], sv[SR];
#pragma unroll
            for (int q = 0; q < SR + 2; q++) {
                const int y = ty0 - M + r0 + q - 1;
                uv[q] = (xok && y >= 0 && y < ng && r0 + q - 1 < RE + 1) ? u[y * ng + x]
                                                                         : make_double2(0, 0);
            }
#pragma unroll
            for (int q = 0; q < SR; q++) sv[q] = sat(x, ty0 - M + r0 + q);
#pragma unroll
            for (int q = 0; q < SR; q++) {
                const int r = r0 + q, y = ty0 - M + r;
                if (r < RE) {
                    double2 v = make_double2(0, 0);
                    if (xok && y >= 0 && y < ng) {
                        const double2 uc = uv[q + 1];
                        const double2 st = cadd(csub(uv[q + 2], cscale(uc, 2.0)), uv[q]);
                        v = cadd(cadd(sv[q], uc), make_double2(-a * st.y, a * st.x));
                    }
                    E[r * PE + c] = v;
                }
            }
        }
    }
    __syncthreads();
    // implicit x: u' = sum_m band[x][m] E(x+m-M, y) on rows [ty0-M, ty0+TS+M);
    // lane -> (row r = t % RE, columns [CX q, CX q + CX)), CX outputs from one
    // register window.  Then the y half step's right-hand side
    // S + (1 + i a d_x^2) u' = S + 2 u' - E  (1 + i a d^2 = 2 - A, A u' = E):
    // no x stencil on u'.
    const bool xact = t < RE * (TS / CX);
    double2 fx[CX];
    if (xact) {
        const int r = t % RE, c0 = (t / RE) * CX;
        const int y = ty0 - M + r;
        double2 ev[CX + 2 * M];
#pragma unroll
        for (int q = 0; q < CX + 2 * M; q++) ev[q] = E[r * PE + c0 + q];
        double2 acc[CX];
#pragma unroll
        for (int j = 0; j < CX; j++) acc[j] = make_double2(0, 0);
        if (xin) {  // uniform: the shared band through the scalar cache, symmetric form
            const double2 b0 = bc[M];
#pragma unroll
            for (int j = 0; j < CX; j++) acc[j] = cmul(b0, ev[j + M]);
#pragma unroll
            for (int k = 1; k <= M; k++) {
                const double2 b = bc[M + k];
#pragma unroll
                for (int j = 0; j < CX; j++) acc[j] = cfma(b, cadd(ev[j + M - k], ev[j + M + k]), acc[j]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < CX; j++)
#pragma unroll
                for (int m = 0; m < W; m++) acc[j] = cfma(BX[(c0 + j) * W + m], ev[j + m], acc[j]);
        }
#pragma unroll
        for (int j = 0; j < CX; j++) {
            const int x = tx0 + c0 + j;
            double2 f = make_double2(0, 0);
            if (x < ng && y >= 0 && y < ng)
                f = csub(cadd(sat(x, y), cscale(acc[j], 2.0)), ev[j + M]);
            fx[j] = f;
        }
    }
    __syncthreads();  // every read of E is done: F and the y band go over it
    if (xact) {
        const int r = t % RE, c0 = (t / RE) * CX;
#pragma unroll
        for (int j = 0; j < CX; j++) G[r * PG + c0 + j] = fx[j];
    }
    if (!yin)
        for (int e = t; e < TS * W; e += NT) {
            const int y = ty0 + e / W;
            BY[e] = y < ng ? band[(long long)y * W + e % W] : make_double2(0, 0);
        }
    __syncthreads();
    // implicit y: u'' = sum_m band[y][m] F(x, y+m-M), four consecutive y per lane
    const double shift = -0.5 * (double)(ng - 1);
    const int icen = (ng / 2) * ng + (ng / 2);
    double dg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    {
        const int c = t % TS, r0 = (t / TS) * 4;
        const int x = tx0 + c;
        double2 fv[4 + 2 * M];
#pragma unroll
        for (int q = 0; q < 4 + 2 * M; q++) fv[q] = G[(r0 + q) * PG + c];
        double2 acc4[4];
#pragma unroll
        for (int j = 0; j < 4; j++) acc4[j] = make_double2(0, 0);
        if (yin) {
            const double2 b0 = bc[M];
#pragma unroll
            for (int j = 0; j < 4; j++) acc4[j] = cmul(b0, fv[j + M]);
#pragma unroll
            for (int k = 1; k <= M; k++) {
                const double2 b = bc[M + k];
#pragma unroll
                for (int j = 0; j < 4; j++) acc4[j] = cfma(b, cadd(fv[j + M - k], fv[j + M + k]), acc4[j]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int m = 0; m < W; m++) acc4[j] = cfma(BY[(r0 + j) * W + m], fv[j + m], acc4[j]);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int y = ty0 + r0 + j;
            const double2 acc = acc4[j];
            if (x < ng && y < ng) {
                uo[y * ng + x] = acc;
                const double wei = acc.x * acc.x + acc.y * acc.y;
                const double dx = (double)x + shift, dy = (double)y + shift;
                dg[0] += wei;
                dg[1] += dx * wei;
                dg[2] += dx * dx * wei;
                dg[3] += dy * wei;
                dg[4] += dy * dy * wei;
                dg[5] += acc.x;
                dg[6] += acc.y;
                if (y * ng + x == icen) dg[7] = wei;
            }
        }
    }
    double *red = (double *)sm;
    block_sum_waves<NT, 8>(dg, red);
    if (t == 0) {
        double *o = tilediag + ((long long)fs * ntiles + tile) * 8;
        for (int k = 0; k < 8; k++) o[k] = dg[k];
    }
}

// Combine the per-tile diagnostic sums of each storage slice in tile order.
__global__ void __launch_bounds__(64) k_adi_tile_diag(const double *__restrict__ tilediag,
                                                      int ntiles, int ns,
                                                      double *__restrict__ diag) {
    const int fs = blockIdx.x * 64 + threadIdx.x;
    if (fs >= ns) return;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < ntiles; k++)
        for (int q = 0; q < 8; q++) v[q] += tilediag[((long long)fs * ntiles + k) * 8 + q];
    for (int q = 0; q < 8; q++) diag[8 * fs + q] = v[q];
}

// ---------------------------------------------------------------------------
// K5+K6 as a whole-width stream: the banded operator of k_adi_band, both
// half steps of one slice per workgroup of NW = blockDim / 64 wavefronts,
// one lane per column, walking down the rows.  Row r:
//   E(r)     = S + u + i a (u(r+1) - 2 u(r) + u(r-1))       (own column)
//   u'(r)    = sum_m band[x][m] E(r, x+m-M)                  (E row via LDS)
//   F(r)     = S + 2 u'(r) - E(r)                             (register ring)
//   u''(r-M) = sum_m band[r-M][m] F(r-2M+m)                   (ring of W rows)
// The products of k_adi_band, each summed as four partial sums (band terms
// k mod 4: independent FMA chains); no cell is computed twice (a
// 32 x 32 tile of k_adi_band evaluates its x stage on 32 + 2M rows and
// loads a halo of M + 1 cells) and F never leaves the registers.
//
// Column map: wavefront 0 takes the first and the last 32 columns, where
// the band rows differ from the central one (their rows sit in LDS); the
// other wavefronts take the columns in between and use the symmetric
// central row (the host offers this kernel when [klo, khi] covers them).
// The E row is double-buffered by row parity with one LDS-only barrier per
// row; the field and source rows are loaded two rows ahead; the F ring has
// W slots with row r in slot r mod W (the row loop is unrolled by W).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void adi_lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// a zero cell the streaming kernel reads for source rows outside the box
// (a select between two global addresses: a local zero would be addressed
// in scratch memory)
__device__ double2 g_zero_cell64;

template <int M>
size_t wide64_lds_bytes(int nw) {
    return (size_t)2 * (64 * nw + 2 * M) * sizeof(double2) + (size_t)(2 * M + 1) * 64 * sizeof(double2) +
           (size_t)nw * 8 * sizeof(double);
}

template <int M>
__global__ void __launch_bounds__(512, 3) k_adi_wide64(const double2 *__restrict__ fin,
                                                     double2 *__restrict__ fout,
                                                     const double2 *__restrict__ src,
                                                     const int4 *__restrict__ box,
                                                     const double2 *__restrict__ band, double a, int ng,
                                                     int first, int ns, int klo, int khi,
                                                     double *__restrict__ diag) {
    constexpr int W = 2 * M + 1;
    extern __shared__ double2 wsm[];
    const int nt = blockDim.x, nw = nt >> 6;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int lp = nt + 2 * M;
    double2 *lrow = wsm;                           // [2][lp] E rows, cell x at x + M
    double2 *lband = wsm + 2 * lp;                 // [W][64] band rows of wavefront 0's columns
    double *wsum = (double *)(lband + W * 64);     // [nw][8]
    const int ii = blockIdx.x;
    if (ii >= ns) return;
    // this lane's column (lanes left over sit on column ng: zero)
    int x;
    if (ng <= 64) x = lane < ng ? lane : ng;
    else if (w == 0) x = lane < 32 ? lane : ng - 64 + lane;
    else {
        x = 32 + 64 * (w - 1) + lane;
        if (x >= ng - 32) x = ng;
    }
    const bool xin = x < ng;
    bool wsym;  // every grid column of this wavefront uses the central band row (uniform)
    if (ng <= 64 || w == 0) wsym = klo <= 0 && khi >= ng - 1;
    else {
        const int xlo = 32 + 64 * (w - 1), xhi = min(xlo + 63, ng - 33);
        wsym = xlo > xhi || (xlo >= klo && xhi <= khi);
    }
    for (int k = t; k < 2 * lp; k += nt) lrow[k] = make_double2(0, 0);
    if (w == 0)
        for (int m = 0; m < W; m++)
            lband[m * 64 + lane] = xin ? band[(long long)x * W + m] : make_double2(0, 0);
    const double2 *bc = band + (long long)((klo + khi) / 2) * W;
    const long long fs = (long long)((ii + first) % ns);
    const double2 *u = fin + fs * ng * ng;
    double2 *uo = fout + fs * ng * ng;
    const int4 bx = box[ii];
    const bool sxin = xin && x >= bx.x && x <= bx.z;
    const int xcl = min(x, ng - 1);
    const double2 *ucol = u + xcl;
    const double2 *scol = src + (long long)ii * ng * ng + xcl;
    const double2 zero = make_double2(0, 0);
    auto uload = [&](int rr) -> double2 { return ucol[(long long)min(rr, ng - 1) * ng]; };
    auto sload = [&](int rr) -> double2 {
        return *((sxin && rr >= bx.y && rr <= bx.w) ? scol + (long long)rr * ng : &g_zero_cell64);
    };
    // u rows r-1 .. r+3 and S rows r .. r+2 at the start of row r
    double2 ul = zero, uc = uload(0), un = uload(1), up1 = uload(2), up2 = uload(3);
    double2 sv = sload(0), sp1 = sload(1), sp2 = sload(2);
    double2 ring[W];
#pragma unroll
    for (int m = 0; m < W; m++) ring[m] = zero;
    // field sums of this column (x fixed: the x moments are dx, dx^2 times
    // the column's power)
    double s_w = 0, s_yw = 0, s_yyw = 0, s_re = 0, s_im = 0, s_cen = 0;
    const bool xcen = x == ng / 2;
    const double shift = -0.5 * (double)(ng - 1);
    adi_lds_barrier();
    double2 *lx = lrow + x + M;
    const int nrow = ng + M;
    for (int r0 = 0; r0 < nrow; r0 += W) {
#pragma unroll
        for (int j = 0; j < W; j++) {
            const int r = r0 + j;
            double2 F = zero;
            if (r < ng) {
                // explicit y stencil of the x half step (FieldSolverADI.cpp:64-72)
                double2 E = zero;
                if (xin) {
                    const double2 a1 = (r + 1 < ng) ? un : zero, a0 = (r > 0) ? ul : zero;
                    const double2 st = cadd(csub(a1, cscale(uc, 2.0)), a0);
                    E = cadd(cadd(sv, uc), make_double2(-a * st.y, a * st.x));
                }
                double2 *lr = lx + (r & 1) * lp;
                lr[0] = E;
                adi_lds_barrier();
                // implicit x and the y half step's right-hand side
                if (xin) {
                    // four partial sums (band terms k mod 4): independent FMA
                    // chains for the latency of the FP64 pipeline
                    double2 pa[4] = {zero, zero, zero, zero};
                    if (wsym) {
                        pa[0] = cmul(bc[M], lr[0]);
#pragma unroll
                        for (int k = 1; k <= M; k++) pa[k & 3] = cfma(bc[M + k], cadd(lr[-k], lr[k]), pa[k & 3]);
                    } else {
#pragma unroll
                        for (int m = 0; m < W; m++) pa[m & 3] = cfma(lband[m * 64 + lane], lr[m - M], pa[m & 3]);
                    }
                    const double2 acc = cadd(cadd(pa[0], pa[1]), cadd(pa[2], pa[3]));
                    F = csub(cadd(sv, cscale(acc, 2.0)), E);
                }
                ul = uc;
                uc = un;
                un = up1;
                up1 = up2;
                up2 = uload(r + 4);
                sv = sp1;
                sp1 = sp2;
                sp2 = sload(r + 3);
            }
            ring[j] = F;  // row r in slot r mod W
            // F rows r-2M .. r (window index k in slot (j+1+k) mod W): a copy
            // in index order, renamed away once the loop is unrolled
            double2 rw[W];
#pragma unroll
            for (int k = 0; k < W; k++) rw[k] = ring[(j + 1 + k) % W];
            const int y = r - M;
            if (y >= 0 && y < ng && xin) {
                // implicit y on F rows y-M .. y+M
                double2 pa[4] = {zero, zero, zero, zero};
                if (y >= klo && y <= khi) {
                    pa[0] = cmul(bc[M], rw[M]);
#pragma unroll
                    for (int k = 1; k <= M; k++) pa[k & 3] = cfma(bc[M + k], cadd(rw[M - k], rw[M + k]), pa[k & 3]);
                } else {
                    const double2 *by = band + (long long)y * W;
#pragma unroll
                    for (int m = 0; m < W; m++) pa[m & 3] = cfma(by[m], rw[m], pa[m & 3]);
                }
                const double2 acc = cadd(cadd(pa[0], pa[1]), cadd(pa[2], pa[3]));
                uo[(long long)y * ng + x] = acc;
                const double wei = acc.x * acc.x + acc.y * acc.y;
                const double dy = (double)y + shift;
                s_w += wei;
                s_yw += dy * wei;
                s_yyw += dy * dy * wei;
                s_re += acc.x;
                s_im += acc.y;
                if (xcen && y == ng / 2) s_cen = wei;
            }
        }
    }
    const double dx = (double)x + shift;
    double dg[8] = {s_w, dx * s_w, dx * dx * s_w, s_yw, s_yyw, s_re, s_im, s_cen};
    if (!xin)
        for (int k = 0; k < 8; k++) dg[k] = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
        for (int o = 32; o > 0; o >>= 1) dg[k] += __shfl_xor(dg[k], o);
    if (lane == 0)
        for (int k = 0; k < 8; k++) wsum[w * 8 + k] = dg[k];
    __syncthreads();
    if (t < 8) {
        double v = wsum[t];
        for (int q = 1; q < nw; q++) v += wsum[q * 8 + t];
        diag[8 * fs + t] = v;
    }
}

// Zero one field slice (slippage inflow, Control.cpp:255-267).  A plain
// store kernel rather than a memset so the step can be graph-captured.
__global__ void __launch_bounds__(256) k_zero(double2 *__restrict__ p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double2(0, 0);
}

// ---------------------------------------------------------------------------
// K7: diagnostics.  Beam part of DiagBeam::getValues (Diagnostic.cpp:387-489):
// out[NB*is + ...] = g1 g2 x1 x2 y1 y2 bre bim (sums; normalised on host).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_diag_beam(const long long *__restrict__ off,
                                                   const double *__restrict__ g,
                                                   const double *__restrict__ th,
                                                   const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   double *__restrict__ out) {
    __shared__ double lds[8 * 256];
    const int is = blockIdx.x;
    const long long a = off[is], b = off[is + 1];
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long long i = a + threadIdx.x; i < b; i += 256) {
        double xi = x[i], yi = y[i], gi = g[i];
        v[0] += gi;
        v[1] += gi * gi;
        v[2] += xi;
        v[3] += xi * xi;
        v[4] += yi;
        v[5] += yi * yi;
        double s, c;
        sincos((0 + 1) * th[i], &s, &c);
        v[6] += c;
        v[7] += s;
    }
    block_sum<256, 8>(v, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; k++) out[8 * is + k] = v[k];
}

// Harmonic bunching sums of DiagBeam::getValues (Diagnostic.cpp:431-436) per
// beam slice for harmonics h0 .. h0 + cnt - 1 (cnt <= 8):
// out[2 (is cnt + k) + {0, 1}] = sum cos((h0 + k) th), sum sin((h0 + k) th).
__global__ void __launch_bounds__(256) k_diag_bunch(const long long *__restrict__ off,
                                                    const double *__restrict__ th, int h0, int cnt,
                                                    double *__restrict__ out) {
    __shared__ double lds[16 * 256];
    const int is = blockIdx.x;
    const long long a = off[is], b = off[is + 1];
    double v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = 0;
    for (long long i = a + threadIdx.x; i < b; i += 256) {
        const double t = th[i];
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (k < cnt) {
                double sn, cs;
                sincos((double)(h0 + k) * t, &sn, &cs);
                v[2 * k] += cs;
                v[2 * k + 1] += sn;
            }
    }
    block_sum<256, 16>(v, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < cnt; k++) {
            out[2 * ((size_t)is * cnt + k)] = v[2 * k];
            out[2 * ((size_t)is * cnt + k) + 1] = v[2 * k + 1];
        }
}

// Field part of DiagField::getValues (:754-861) per storage slice:
// out[8*is0 + ...] = power x1 x2 y1 y2 ffre ffim inten(on-axis |E|^2).
// Auxiliary beam diagnostics of DiagBeam::getValues (Diagnostic.cpp:415-445,
// 505-519) per beam slice: sums of px and py, the extremes of x, px, y, py and
// gamma (starting from the reference's bounds 1e5 / -1e5, gamma 1e7 / 1),
// and the on-axis value of the slice's field (DiagField, :855-858).
// out[16 is + ..] = {sum px, sum py, xmin, xmax, pxmin, pxmax, ymin, ymax,
// pymin, pymax, gmin, gmax, Re E0, Im E0}.
__device__ __forceinline__ void minmax_block(double (&m)[10], double *lds) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
#pragma unroll
    for (int k = 0; k < 10; k++)
        for (int o = 32; o > 0; o >>= 1) {
            const double u = __shfl_xor(m[k], o);
            m[k] = (k & 1) ? fmax(m[k], u) : fmin(m[k], u);
        }
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 10; k++) lds[wv * 10 + k] = m[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 10; k++) {
        double v = lds[k];
        for (int q = 1; q < (int)(blockDim.x >> 6); q++)
            v = (k & 1) ? fmax(v, lds[q * 10 + k]) : fmin(v, lds[q * 10 + k]);
        m[k] = v;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_diag_aux(const long long *__restrict__ off,
                                                  const double *__restrict__ g,
                                                  const double *__restrict__ x,
                                                  const double *__restrict__ y,
                                                  const double *__restrict__ px,
                                                  const double *__restrict__ py,
                                                  const double2 *__restrict__ field, int ng, int ns,
                                                  int first, double *__restrict__ out) {
    __shared__ double lds[10 * 256];
    const int is = blockIdx.x;
    const long long a = off[is], b = off[is + 1];
    double v[2] = {0, 0};
    double m[10] = {1e5, -1e5, 1e5, -1e5, 1e5, -1e5, 1e5, -1e5, 1e7, 1};
    for (long long i = a + threadIdx.x; i < b; i += 256) {
        const double xi = x[i], yi = y[i], pxi = px[i], pyi = py[i], gi = g[i];
        v[0] += pxi;
        v[1] += pyi;
        m[0] = fmin(m[0], xi);
        m[1] = fmax(m[1], xi);
        m[2] = fmin(m[2], pxi);
        m[3] = fmax(m[3], pxi);
        m[4] = fmin(m[4], yi);
        m[5] = fmax(m[5], yi);
        m[6] = fmin(m[6], pyi);
        m[7] = fmax(m[7], pyi);
        m[8] = fmin(m[8], gi);
        m[9] = fmax(m[9], gi);
    }
    block_sum<256, 2>(v, lds);
    minmax_block(m, lds);
    if (threadIdx.x == 0) {
        double *o = out + 16 * (long long)is;
        o[0] = v[0];
        o[1] = v[1];
        for (int k = 0; k < 10; k++) o[2 + k] = m[k];
        const double2 c = field[(long long)((is + first) % ns) * ng * ng + (ng / 2) * ng + ng / 2];
        o[12] = c.x;
        o[13] = c.y;
    }
}

__global__ void __launch_bounds__(256) k_diag_field(const double2 *__restrict__ field, int ng,
                                                    double *__restrict__ out) {
    __shared__ double lds[8 * 256];
    const int is0 = blockIdx.x;
    const double2 *f = field + (long long)is0 * ng * ng;
    const double shift = -0.5 * (double)(ng - 1);
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < ng * ng; i += 256) {
        int iy = i / ng, ix = i % ng;
        double dy = (double)iy + shift, dx = (double)ix + shift;
        double2 loc = f[i];
        double wei = loc.x * loc.x + loc.y * loc.y;
        v[0] += wei;
        v[1] += dx * wei;
        v[2] += dx * dx * wei;
        v[3] += dy * wei;
        v[4] += dy * dy * wei;
        v[5] += loc.x;
        v[6] += loc.y;
    }
    block_sum<256, 8>(v, lds);
    if (threadIdx.x == 0) {
        int ic = (ng / 2) * ng + (ng / 2);
        double2 c = f[ic];
        v[7] = c.x * c.x + c.y * c.y;
        for (int k = 0; k < 8; k++) out[8 * is0 + k] = v[k];
    }
}

// Sorting::localSort target slice (src/Util/Sorting.cpp:96).
__global__ void __launch_bounds__(256) k_sort_index(long long n, const double *__restrict__ th,
                                                    double invslen, int *__restrict__ atar,
                                                    unsigned long long *__restrict__ moved) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int m = 0;
    if (i < n) {
        int at = (int)floor(th[i] * invslen);
        atar[i] = at;
        m = (at != 0);
    }
    unsigned long long bal = __ballot(m);
    if ((threadIdx.x & 63) == 0 && bal) atomicAdd(moved, (unsigned long long)__popcll(bal));
}


// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void launch_transverse(hipStream_t s, long long n, const double *g, double *x, double *y,
                       double *px, double *py, const TransParams &p) {
    if (n > 0) hipLaunchKernelGGL(k_transverse, dim3(nblk(n, 256)), dim3(256), 0, s, n, g, x, y, px, py, p);
}
__global__ void __launch_bounds__(256) k_kicks(const long long *__restrict__ off, double *__restrict__ g,
                                                const double *__restrict__ bdg, const long long *__restrict__ boff,
                                                int nbins, const double *__restrict__ sdg) {
    const int s = blockIdx.x;
    const long long a = off[s], b = off[s + 1];
    const double ds = sdg ? sdg[s] : 0.0;
    const long long b0 = bdg ? boff[s] : 0;
    for (long long i = a + threadIdx.x; i < b; i += blockDim.x) {
        double v = g[i];
        if (bdg) v += bdg[b0 + (i - a) / nbins];
        if (sdg) v += ds;
        g[i] = v;
    }
}

void launch_kicks(hipStream_t s, int ns, const long long *off, double *g, const double *beamlet_dg,
                  const long long *boff, int nbins, const double *slice_dg) {
    if (ns > 0 && (beamlet_dg || slice_dg))
        hipLaunchKernelGGL(k_kicks, dim3(ns), dim3(256), 0, s, off, g, beamlet_dg, boff, nbins, slice_dg);
}

void launch_r56(hipStream_t s, long long n, const double *g, double *th, double r56, double g0) {
    if (n > 0) hipLaunchKernelGGL(k_r56, dim3(nblk(n, 256)), dim3(256), 0, s, n, g, th, r56, g0);
}
void launch_slice_moments(hipStream_t s, int ns, const long long *off, const double *x,
                          const double *y, double *mom) {
    hipLaunchKernelGGL(k_slice_moments, dim3(ns), dim3(256), 0, s, off, x, y, mom);
}
void launch_slice_sizes(hipStream_t s, int ns, const long long *off, const double *cur,
                        const double *mom, double *fsize, double *cen) {
    hipLaunchKernelGGL(k_slice_sizes, dim3(nblk(ns, 256)), dim3(256), 0, s, ns, off, cur, mom, fsize, cen);
}
void launch_lsc_long(hipStream_t s, int ns, const double *cur, const double *fsize, double sl,
                     double gamma, double scl, double *esc) {
    hipLaunchKernelGGL(k_lsc_long, dim3(ns), dim3(64), 0, s, ns, cur, fsize, sl, gamma, scl, esc);
}
void launch_sc_rbound(hipStream_t s, int ns, const long long *off, const double *x,
                      const double *y, const double *cen, double *rbound) {
    hipLaunchKernelGGL(k_sc_rbound, dim3(ns), dim3(256), 0, s, off, x, y, cen, rbound);
}
void launch_sc_scan(hipStream_t s, int ns, const long long *off, const double *rbound,
                    double *rmax_state, int ngr, double *dr) {
    hipLaunchKernelGGL(k_sc_scan, dim3(1), dim3(kScanNT), 0, s, ns, off, rbound, rmax_state, ngr, dr);
}
void launch_sc_solve(hipStream_t s, int ns, const long long *off, const double *x,
                     const double *y, const double *th, const double *cur, const double *cen,
                     const double *dr, int ngr, int nz, int nphi, double ks, double gz2,
                     double *ez, double *efield) {
    hipLaunchKernelGGL(k_sc_solve, dim3(ns), dim3(256), 0, s, off, x, y, th, cur, cen, dr, ngr, nz,
                       nphi, ks, gz2, ez, efield);
}
void launch_push(hipStream_t s, int nblocks, const int *blk_slice, const long long *blk_begin,
                 const long long *off, double *g, double *th, const double *x, const double *y,
                 const double *px, const double *py, const double2 *field, const double *esc,
                 const double *ezsc, const PushParams &p) {
    if (nblocks <= 0) return;
#define G4_PUSH(NH)                                                                                     \
    hipLaunchKernelGGL(k_push<NH>, dim3(nblocks), dim3(256), 0, s, blk_slice, blk_begin, off, g, th, x, y, \
                       px, py, field, esc, ezsc, p)
    switch (p.nh) {
        case 1: G4_PUSH(1); break;
        case 2: G4_PUSH(2); break;
        case 3: G4_PUSH(3); break;
        case 4: G4_PUSH(4); break;
        default: G4_PUSH(0); break;
    }
#undef G4_PUSH
}
void launch_fp64_front(hipStream_t s, int ns, const long long *off, const double *g, double *x, double *y,
                       double *px, double *py, const double *cur, const TransParams &t1, bool moments, bool sc,
                       double gridmax, double dgrid, int ng, double *mom, double *fsize, double *cen,
                       double *rbound, int4 *pbox, const DepositParams &d, double *pmaxf) {
    hipLaunchKernelGGL(k_fp64_front, dim3(ns), dim3(kFrontNT), 0, s, off, g, x, y, px, py, cur, t1, moments ? 1 : 0,
                       sc ? 1 : 0, gridmax, dgrid, ng, mom, fsize, cen, rbound, pbox, d, pmaxf);
}

int fp64_main_modes() { return kFuseModes; }

void launch_fp64_main(hipStream_t s, int ns, const long long *off, double *g, double *th, double *x, double *y,
                      double *px, double *py, const double2 *field, const double *esc, const double *cur,
                      const double *cen, const double *dr, double *efield, const int4 *pbox,
                      const double *pmaxf, double2 *src, int4 *box, double *beamsum, const FuseParams &f) {
    static bool attr = false;
    if (!attr) {  // 64 KiB of dynamic LDS plus the static reduction scratch
        (void)hipFuncSetAttribute((const void *)k_fp64_main, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  kFuseDyn * sizeof(double2));
        attr = true;
    }
    hipLaunchKernelGGL(k_fp64_main, dim3(ns), dim3(kFuseNT), kFuseDyn * sizeof(double2), s, off, g, th, x, y, px,
                       py, field, esc, cur, cen, dr, efield, pbox, pmaxf, src, box, beamsum, f);
}

void launch_deposit(hipStream_t s, int ns, const long long *off, const double *g,
                    const double *th, const double *x, const double *y, const double *cur,
                    double2 *src, int4 *box, const DepositParams &p) {
    static bool attr = false;
    if (!attr) {  // the LDS tile exceeds the 64 KiB default dynamic limit
        (void)hipFuncSetAttribute((const void *)k_deposit,
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  kDepTile * sizeof(double2));
        attr = true;
    }
    hipLaunchKernelGGL(k_deposit, dim3(ns), dim3(1024), kDepTile * sizeof(double2), s, off, g, th,
                       x, y, cur, src, box, p);
}
void launch_adi(hipStream_t s, int ns, double2 *field, double2 *work, const double2 *src,
                const int4 *box, const AdiFactors &f, int first) {
    dim3 grid(nblk(f.ngrid, 64), ns);
    hipLaunchKernelGGL(k_adi_x, grid, dim3(64), 0, s, field, work, src, box, f, first, ns);
    hipLaunchKernelGGL(k_adi_y, grid, dim3(64), 0, s, field, work, src, box, f, first, ns);
}
template <int T, int G, bool PREF>
static void launch_adi_fused_t(hipStream_t s, int ns, double2 *field, double2 *work,
                               const double2 *src, const int4 *box, const AdiFactors &f,
                               int first, double *diag) {
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_adi_fused<T, G, PREF>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    constexpr int NT = G <= 192 ? 192 : 256;
    const size_t lds = (size_t)(2 * f.ngrid * (T + 1) + 3 * f.ngrid) * sizeof(double2);
    hipLaunchKernelGGL((k_adi_fused<T, G, PREF>), dim3(ns), dim3(NT), lds, s, field, work, src,
                       box, f, first, ns, diag);
}

int launch_adi_fused(hipStream_t s, int ns, double2 *field, double2 *work, const double2 *src,
                     const int4 *box, const AdiFactors &f, int first, double *diag, int tile,
                     int prefetch) {
    if (f.ngrid > 256 || (tile != 8 && tile != 16)) return -1;
    const bool small = f.ngrid <= 192;
#define G4_ADI(T, G, P) launch_adi_fused_t<T, G, P>(s, ns, field, work, src, box, f, first, diag)
    if (tile == 8) {
        if (prefetch) { if (small) G4_ADI(8, 192, true); else G4_ADI(8, 256, true); }
        else { if (small) G4_ADI(8, 192, false); else G4_ADI(8, 256, false); }
    } else {
        if (prefetch) { if (small) G4_ADI(16, 192, true); else G4_ADI(16, 256, true); }
        else { if (small) G4_ADI(16, 192, false); else G4_ADI(16, 256, false); }
    }
#undef G4_ADI
    return 0;
}
template <int M>
static void launch_adi_band_t(hipStream_t s, int ns, const double2 *fin, double2 *fout,
                              const double2 *src, const int4 *box, const double2 *band, double a,
                              int ng, int first, int klo, int khi, double *tilediag,
                              double *diag) {
    constexpr int TS = BandGeom<M>::TS;
    static bool attr = false;
    const size_t lds = BandGeom<M>::lds_cells * sizeof(double2);
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_adi_band<M>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr = true;
    }
    const int ntx = (ng + TS - 1) / TS;
    const unsigned nblocks = (unsigned)(((ns + 7) / 8) * 8) * (unsigned)(ntx * ntx);
    hipLaunchKernelGGL(k_adi_band<M>, dim3(nblocks), dim3(256), lds, s, fin, fout, src, box,
                       band, a, ng, first, ns, ntx, klo, khi, tilediag);
    hipLaunchKernelGGL(k_adi_tile_diag, dim3((ns + 63) / 64), dim3(64), 0, s, tilediag,
                       ntx * ntx, ns, diag);
}

int adi_band_tiles(int ng) {
    const int ntx = (ng + 31) / 32;
    return ntx * ntx;
}

bool adi_wide64_ok(int ng, int m, int klo, int khi) {
    return (m == 7 || m == 8) && ng <= 512 && (ng <= 64 || (klo <= 32 && khi >= ng - 33));
}

template <int M>
static void launch_adi_wide64_t(hipStream_t s, int ns, const double2 *fin, double2 *fout, const double2 *src,
                                const int4 *box, const double2 *band, double a, int ng, int first, int klo,
                                int khi, double *diag) {
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_adi_wide64<M>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)wide64_lds_bytes<M>(8));
        attr = true;
    }
    const int nw = std::max(1, (ng + 63) / 64);
    hipLaunchKernelGGL(k_adi_wide64<M>, dim3(ns), dim3(64 * nw), wide64_lds_bytes<M>(nw), s, fin, fout, src, box,
                       band, a, ng, first, ns, klo, khi, diag);
}

int launch_adi_wide64(hipStream_t s, int ns, const double2 *fin, double2 *fout, const double2 *src,
                      const int4 *box, const double2 *band, double a, int ng, int first, int m, int klo, int khi,
                      double *diag) {
    if (!adi_wide64_ok(ng, m, klo, khi)) return -1;
    if (m == 7) launch_adi_wide64_t<7>(s, ns, fin, fout, src, box, band, a, ng, first, klo, khi, diag);
    else launch_adi_wide64_t<8>(s, ns, fin, fout, src, box, band, a, ng, first, klo, khi, diag);
    return 0;
}

int launch_adi_band(hipStream_t s, int ns, const double2 *fin, double2 *fout, const double2 *src,
                    const int4 *box, const double2 *band, double a, int ng, int first, int m,
                    int klo, int khi, double *tilediag, double *diag) {
#define G4_BAND(MM) launch_adi_band_t<MM>(s, ns, fin, fout, src, box, band, a, ng, first, klo, khi, tilediag, diag)
    if (m == 7) G4_BAND(7);
    else if (m == 8) G4_BAND(8);
    else if (m == 12) G4_BAND(12);
    else if (m == 16) G4_BAND(16);
    else return -1;
#undef G4_BAND
    return 0;
}

// Copy n doubles into pinned host memory (its device address) with plain
// stores on the compute queue: a record then costs no copy-engine transfer
// and no hand-off between the copy and compute queues.
__global__ void __launch_bounds__(256) k_copy_out(const double *__restrict__ src, double *__restrict__ dst,
                                                  long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}
void launch_copy_out(hipStream_t s, const double *src, double *dst, long long n) {
    if (n > 0) hipLaunchKernelGGL(k_copy_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
}
void launch_zero(hipStream_t s, double2 *p, long long n) {
    if (n > 0) hipLaunchKernelGGL(k_zero, dim3(nblk(n, 256)), dim3(256), 0, s, p, n);
}
void launch_diag_beam(hipStream_t s, int ns, const long long *off, const double *g,
                      const double *th, const double *x, const double *y, double *out) {
    hipLaunchKernelGGL(k_diag_beam, dim3(ns), dim3(256), 0, s, off, g, th, x, y, out);
}
void launch_diag_aux(hipStream_t s, int ns, const long long *off, const double *g, const double *x,
                     const double *y, const double *px, const double *py, const double2 *field, int ng,
                     int first, double *out) {
    hipLaunchKernelGGL(k_diag_aux, dim3(ns), dim3(256), 0, s, off, g, x, y, px, py, field, ng, ns, first, out);
}
void launch_diag_bunch(hipStream_t s, int ns, const long long *off, const double *th, int h0, int cnt,
                       double *out) {
    if (ns > 0 && cnt > 0) hipLaunchKernelGGL(k_diag_bunch, dim3(ns), dim3(256), 0, s, off, th, h0, cnt, out);
}
void launch_diag_field(hipStream_t s, int ns, const double2 *field, int ng, double *out) {
    hipLaunchKernelGGL(k_diag_field, dim3(ns), dim3(256), 0, s, field, ng, out);
}
void launch_sort_index(hipStream_t s, long long n, const double *th, double invslen, int *atar,
                       unsigned long long *moved) {
    if (n > 0) hipLaunchKernelGGL(k_sort_index, dim3(nblk(n, 256)), dim3(256), 0, s, n, th, invslen, atar, moved);
}

}  // namespace g4mi
