// FP64 device kernels of the undulator step (plain form).
//
// This translation unit is compiled with -ffp-contract=off so every FP64
// expression rounds exactly as the reference's (non-FMA) CPU build: with
// identical inputs, + - * / sqrt agree bit for bit; sin/cos/log/exp/hypot
// may differ from glibc by an ulp.  Each kernel cites the reference loop it
// restates; expressions keep the reference's association order.
#include "g4mi_kernels.h"

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

__global__ void __launch_bounds__(256) k_transverse(long long n, const double *__restrict__ g,
                                                    double *__restrict__ x, double *__restrict__ y,
                                                    double *__restrict__ px,
                                                    double *__restrict__ py, TransParams p) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double gi = g[i], xi = x[i], yi = y[i], pxi = px[i], pyi = py[i];
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
// The j loop runs in the reference's order, so the sum is bit-identical for
// identical sizes.
__global__ void __launch_bounds__(256) k_lsc_long(int ns, const double *__restrict__ cur,
                                                  const double *__restrict__ fsize, double sl,
                                                  double gamma, double scl,
                                                  double *__restrict__ esc) {
    __shared__ double sc[256], ss[256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    for (int j0 = 0; j0 < ns; j0 += 256) {
        __syncthreads();
        if (j0 + threadIdx.x < ns) {
            sc[threadIdx.x] = cur[j0 + threadIdx.x];
            ss[threadIdx.x] = fsize[j0 + threadIdx.x];
        }
        __syncthreads();
        int jn = min(256, ns - j0);
        if (i < ns) {
            for (int jj = 0; jj < jn; jj++) {
                int j = j0 + jj;
                double ds = (double)(j - i) * sl * gamma;
                double coef = 1 - sqrt(ds * ds / (ds * ds + ss[jj]));
                double sgn = (j > i) ? -1. : ((j < i) ? 1. : 0.);
                e += sgn * coef * scl * sc[jj] / ss[jj];
            }
        }
    }
    if (i < ns) esc[i] = e;
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

// The growing radial grid of the short-range solver: rmax is solver state
// that grows to 1.5 rbound whenever a slice exceeds it, visited in slice
// order (EFieldSolver.cpp:159-165).  One lane; ns is at most a few
// thousand.
__global__ void k_sc_scan(int ns, const long long *__restrict__ off,
                          const double *__restrict__ rbound, double *__restrict__ rmax_state,
                          int ngr, double *__restrict__ dr) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double rmax = rmax_state[0];
    for (int is = 0; is < ns; is++) {
        if (off[is + 1] - off[is] >= 1) {
            if (rbound[is] > rmax) rmax = rbound[is] * 1.5;
        }
        dr[is] = rmax / (double)(ngr - 1);
    }
    rmax_state[0] = rmax;
}

// K3b: short-range space charge of one slice per workgroup,
// EFieldSolver::shortRange (:197-258) with constructLaplaceOperator
// (:176-190) and tridiag (:260-272).  Radial charge histograms are private
// per wavefront and summed in wave order, so the result is deterministic.
constexpr int kScMaxNg = 256;
__global__ void __launch_bounds__(256) k_sc_solve(
    const long long *__restrict__ off, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ th, const double *__restrict__ cur, const double *__restrict__ cen,
    const double *__restrict__ dr_arr, int ngr, int nz, int nphi, double ks, double gz2,
    double *__restrict__ ez, double *__restrict__ efield) {
    __shared__ double2 hist[4][kScMaxNg];
    __shared__ double2 celm[kScMaxNg], gam[kScMaxNg];
    __shared__ double vol[kScMaxNg], ldig[kScMaxNg + 1], rlog[kScMaxNg];
    const int is = blockIdx.x, t = threadIdx.x, w = t >> 6;
    const long long a = off[is], b = off[is + 1];
    const int npart = (int)(b - a);
    for (long long i = a + t; i < b; i += 256) ez[i] = 0;
    if (t == 0) efield[is] = 0;
    if (npart < 1) return;
    const double dr = dr_arr[is], xc = cen[2 * is], yc = cen[2 * is + 1];
    const double pi = 2 * asin(1.0);
    if (t == 0) {
        vol[0] = pi * dr * dr;
        rlog[0] = 0.5;
        ldig[0] = 0;
        for (int j = 1; j < ngr; j++) {
            vol[j] = pi * dr * dr * (2 * j + 1);
            ldig[j] = 2 * pi * j;
            rlog[j] = log((double)(j + 1) / (double)j);
        }
        ldig[ngr] = 0;
    }
    const double coef = -gz2 / ks / ks;
    const double econst = kVacimp / kEev * cur[is] / ks / (double)npart;
    for (int m = -nphi; m <= nphi; m++) {
        for (int l = 1; l <= nz; l++) {
            __syncthreads();
            for (int k = t; k < 4 * kScMaxNg; k += 256) hist[k / kScMaxNg][k % kScMaxNg] = make_double2(0, 0);
            __syncthreads();
            for (long long i = a + t; i < b; i += 256) {
                double tx = x[i] - xc, ty = y[i] - yc;
                double radi = sqrt(tx * tx + ty * ty);
                double2 cw = make_double2(tx / radi, ty / radi);
                int bin = (int)floor(radi / dr);
                bin = (bin >= 0 && bin < ngr) ? bin : (bin >= ngr ? ngr - 1 : 0);
                double s, c;
                sincos(l * th[i], &s, &c);
                double2 v = cmul(cpow_int(cw, -m), make_double2(c, -s));
                unsafeAtomicAdd(&hist[w][bin].x, v.x);
                unsafeAtomicAdd(&hist[w][bin].y, v.y);
            }
            __syncthreads();
            if (t == 0) {
                // lmid for this m (:221-223), coefficients (:240-245), tridiag
                auto lmid = [&](int i) {
                    double v = -ldig[i] - ldig[i + 1] - 2. * pi * m * m * rlog[i];
                    if (i == ngr - 1) v -= 2 * pi * ngr;
                    return v;
                };
                auto src = [&](int i) {
                    double2 cs = cadd(cadd(cadd(hist[0][i], hist[1][i]), hist[2][i]), hist[3][i]);
                    return cmul(cs, make_double2(0, econst / l / vol[i]));
                };
                double2 bet;
                double cmid0 = 1 + coef * lmid(0) / l / l / vol[0];
                bet = make_double2(cmid0, 0);
                celm[0] = cdiv(src(0), bet);
                for (int i = 1; i < ngr; i++) {
                    double cupp = coef * ldig[i] / l / l / vol[i - 1];
                    double clow = coef * ldig[i] / l / l / vol[i];
                    double cmid = 1 + coef * lmid(i) / l / l / vol[i];
                    gam[i] = cdiv(make_double2(cupp, 0), bet);
                    bet = csub(make_double2(cmid, 0), cmul(make_double2(clow, 0), gam[i]));
                    celm[i] = cdiv(csub(src(i), cmul(make_double2(clow, 0), celm[i - 1])), bet);
                }
                for (int i = ngr - 2; i > -1; i--) celm[i] = csub(celm[i], cmul(gam[i + 1], celm[i + 1]));
                if (l == 1) efield[is] = hypot(celm[0].x, celm[0].y);
            }
            __syncthreads();
            for (long long i = a + t; i < b; i += 256) {
                double tx = x[i] - xc, ty = y[i] - yc;
                double radi = sqrt(tx * tx + ty * ty);
                double2 cw = make_double2(tx / radi, ty / radi);
                int bin = (int)floor(radi / dr);
                bin = (bin >= 0 && bin < ngr) ? bin : (bin >= ngr ? ngr - 1 : 0);
                double s, c;
                sincos(l * th[i], &s, &c);
                double2 ct = cmul(cpow_int(cw, m), make_double2(c, s));
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
struct Ode {
    double xks, xku, btpar, ez;
    double2 rpart;
    double k2gg, k2pp;
    __device__ __forceinline__ void eval(double tgam, double tthet) {
        double ztemp1 = -2. / xks;
        double s, c;
        sincos(1.0 * tthet, &s, &c);
        double2 ctmp = cmul(rpart, make_double2(c, -s));
        double btper0 = btpar + ztemp1 * ctmp.x;
        double btpar0 = sqrt(1. - btper0 / (tgam * tgam));
        k2pp += xks * (1. - 1. / btpar0) + xku;
        k2gg += ctmp.y / btpar0 / tgam - ez;
    }
};

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
    Ode o;
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
// tile with FP64 LDS atomics, in row bands when it exceeds the tile, and
// written to the slice's source plane.  box[is] = {ix0, iy0, ix1, iy1}
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
    extern __shared__ double2 tile[];
    __shared__ int red[4][1024];
    const int is = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
    const long long a = off[is], b = off[is + 1];
    const int ng = p.ngrid;
    int bx0 = 1 << 30, by0 = 1 << 30, bx1 = -1, by1 = -1;
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
            }
        }
    }
    red[0][t] = bx0;
    red[1][t] = by0;
    red[2][t] = -bx1;
    red[3][t] = -by1;
    __syncthreads();
    for (int s = nt / 2; s > 0; s >>= 1) {
        if (t < s)
            for (int k = 0; k < 4; k++) red[k][t] = min(red[k][t], red[k][t + s]);
        __syncthreads();
    }
    bx0 = red[0][0];
    by0 = red[1][0];
    bx1 = -red[2][0];
    by1 = -red[3][0];
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
    double2 *plane = src + (long long)is * ng * ng;
    for (int r0 = by0; r0 <= by1; r0 += rows_per_band) {
        const int r1 = min(by1, r0 + rows_per_band - 1);
        const int ncell = (r1 - r0 + 1) * ww;
        for (int k = t; k < ncell; k += nt) tile[k] = make_double2(0, 0);
        __syncthreads();
        for (long long i = a + t; i < b; i += nt) {
            double xi = x[i], yi = y[i];
            if (!((xi > -p.gridmax) && (xi < p.gridmax) && (yi > -p.gridmax) && (yi < p.gridmax)))
                continue;
            double wx = (xi + p.gridmax) / p.dgrid;
            double wy = (yi + p.gridmax) / p.dgrid;
            int ix = min((int)floor(wx), ng - 2), iy = min((int)floor(wy), ng - 2);
            if (iy + 1 < r0 || iy > r1) continue;
            wx = 1 + floor(wx) - wx;
            wy = 1 + floor(wy) - wy;
            double ddx = xi - p.ax, ddy = yi - p.ay;
            double faw2 = (1 + p.kx * ddx * ddx + p.ky * ddy * ddy + 2 * (p.gradx * ddx + p.grady * ddy));
            double part = sqrt(faw2) * scl / g[i];
            double s, c;
            sincos(1.0 * th[i], &s, &c);
            double2 cp = make_double2(s * part, c * part);
            double wgt[4] = {wx * wy, (1 - wx) * wy, wx * (1 - wy), (1 - wx) * (1 - wy)};
            int cx[4] = {ix, ix + 1, ix, ix + 1};
            int cy[4] = {iy, iy, iy + 1, iy + 1};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (cy[k] < r0 || cy[k] > r1 || cx[k] > bx1) continue;
                int lk = (cy[k] - r0) * ww + (cx[k] - bx0);
                unsafeAtomicAdd(&tile[lk].x, wgt[k] * cp.x);
                unsafeAtomicAdd(&tile[lk].y, wgt[k] * cp.y);
            }
        }
        __syncthreads();
        for (int k = t; k < ncell; k += nt) {
            int ry = r0 + k / ww, rx = bx0 + k % ww;
            plane[ry * ng + rx] = tile[k];
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

// Field part of DiagField::getValues (:754-861) per storage slice:
// out[8*is0 + ...] = power x1 x2 y1 y2 ffre ffim inten(on-axis |E|^2).
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
    hipLaunchKernelGGL(k_lsc_long, dim3(nblk(ns, 256)), dim3(256), 0, s, ns, cur, fsize, sl, gamma, scl, esc);
}
void launch_sc_rbound(hipStream_t s, int ns, const long long *off, const double *x,
                      const double *y, const double *cen, double *rbound) {
    hipLaunchKernelGGL(k_sc_rbound, dim3(ns), dim3(256), 0, s, off, x, y, cen, rbound);
}
void launch_sc_scan(hipStream_t s, int ns, const long long *off, const double *rbound,
                    double *rmax_state, int ngr, double *dr) {
    hipLaunchKernelGGL(k_sc_scan, dim3(1), dim3(64), 0, s, ns, off, rbound, rmax_state, ngr, dr);
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
    if (nblocks > 0)
        hipLaunchKernelGGL(k_push, dim3(nblocks), dim3(256), 0, s, blk_slice, blk_begin, off, g, th,
                           x, y, px, py, field, esc, ezsc, p);
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
void launch_zero(hipStream_t s, double2 *p, long long n) {
    if (n > 0) hipLaunchKernelGGL(k_zero, dim3(nblk(n, 256)), dim3(256), 0, s, p, n);
}
void launch_diag_beam(hipStream_t s, int ns, const long long *off, const double *g,
                      const double *th, const double *x, const double *y, double *out) {
    hipLaunchKernelGGL(k_diag_beam, dim3(ns), dim3(256), 0, s, off, g, th, x, y, out);
}
void launch_diag_field(hipStream_t s, int ns, const double2 *field, int ng, double *out) {
    hipLaunchKernelGGL(k_diag_field, dim3(ns), dim3(256), 0, s, field, ng, out);
}
void launch_sort_index(hipStream_t s, long long n, const double *th, double invslen, int *atar,
                       unsigned long long *moved) {
    if (n > 0) hipLaunchKernelGGL(k_sort_index, dim3(nblk(n, 256)), dim3(256), 0, s, n, th, invslen, atar, moved);
}

}  // namespace g4mi
