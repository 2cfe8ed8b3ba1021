// Mixed-precision kernels of the undulator step: FP32 particles and field,
// FP64 slice phase base, FP64 residual correction in the ADI and FP64
// accumulation for every diagnostic sum.
//
// Each kernel restates the same reference loop as its FP64 counterpart in
// g4mi_kernels.hip (cited there); the differences are the storage type and
// the offset form of gamma and theta:
//   gamma = gref + dg,  theta = base[slice] + dth.
// In the longitudinal equation (BeamSolver::ODE, src/Core/BeamSolver.cpp:
// 144-163) the phase rate xks (1 - 1/beta_par) + xku is split into the
// reference particle's rate kappa (applied to the FP64 base) and the
// particle's departure
//   dth' = xks (g(eps_ref) - g(eps)),  g(e) = 1/sqrt(1-e) - 1,  eps = btper0/gamma^2,
// evaluated as xks g'(eps_bar) (eps_ref - eps) with eps_ref - eps formed
// from differences that are small in their own right (dg, px^2 + py^2,
// aw^2 (awloc^2 - 1), the field term): no catastrophic cancellation in FP32.
//
// Per-step parameters come from the device step table steps[st->istep], so
// a captured graph replays any step.
#include <algorithm>
#include <type_traits>

#include <hip/hip_cooperative_groups.h>

#include "g4mi_mixed.h"
#include "g4mi_scan.h"

namespace g4mi {

namespace cg = cooperative_groups;

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kPi = 3.14159265358979323846264338328;

__device__ __forceinline__ float2 f2(float a, float b) { return make_float2(a, b); }
__device__ __forceinline__ float2 fadd(float2 a, float2 b) { return f2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 fsub(float2 a, float2 b) { return f2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 fscale(float2 a, float s) { return f2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 fmul(float2 a, float2 b) {
    return f2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 ffma(float2 a, float2 b, float2 acc) {
    acc.x = fmaf(a.x, b.x, fmaf(-a.y, b.y, acc.x));
    acc.y = fmaf(a.x, b.y, fmaf(a.y, b.x, acc.y));
    return acc;
}
__device__ __forceinline__ double2 d2(double a, double b) { return make_double2(a, b); }
// The hardware's FP32 reciprocal, square root and reciprocal square root
// (v_rcp_f32, v_sqrt_f32, v_rsq_f32, about 1 ulp on normal inputs) in place
// of the correctly rounded sequences (a division is ten instructions): the
// mixed form's error is set by its FP32 storage, not by the last ulp here.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double2 tod(float2 a) { return d2((double)a.x, (double)a.y); }
__device__ __forceinline__ float2 tof(double2 a) { return f2((float)a.x, (float)a.y); }
__device__ __forceinline__ float2 tof(float2 a) { return a; }

// phase in [0, 2 pi) of an FP64 angle, as FP32
__device__ __forceinline__ float wrap_phase(double p) {
    return (float)(p - kTwoPi * floor(p / kTwoPi));
}

// Deterministic block sum of K doubles over NT threads (fixed tree order);
// lds holds K*NT doubles.
template <int NT, int K>
__device__ __forceinline__ void bsum(double (&v)[K], double *lds) {
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

// transverse half step (TrackBeam::track, src/Core/TrackBeam.cpp:8-116), FP32.
// The focusing angle per step is small (omg ~ 1e-3), so cos(omg) is 1 to
// within a few FP32 ulp; the map is applied as x + (a1 - 1) x + ... with
// a1 - 1 = -2 sin^2(omg/2) (cosh: +2 sinh^2(omg/2)), which keeps the
// focusing term's full relative precision.
// sinh for the defocusing map: its Taylor series through x^9 below 0.5
// (truncation below 1e-10 relative), exp above
__device__ __forceinline__ float sinh32(float v) {
    const float av = fabsf(v);
    if (av < 0.5f) {
        const float v2 = v * v;
        return v * (1.f + v2 * (1.f / 6.f + v2 * (1.f / 120.f + v2 * (1.f / 5040.f + v2 * (1.f / 362880.f)))));
    }
    const float e = __expf(av), r = 0.5f * (e - frcp(e));
    return v < 0.f ? -r : r;
}

__device__ __forceinline__ void line_map32(int kind, float delz, float q, float &x, float &px,
                                           float gz, float d) {
    // igz = 1 / gz
    const float igz = frcp(gz);
    if (kind == 0) {
        x += px * delz * igz;
    } else if (kind > 0) {
        const float foc = fsqrt(q * igz), omg = foc * delz;
        float s, c, sh, ch;
        __sincosf(omg, &s, &c);
        __sincosf(0.5f * omg, &sh, &ch);
        const float a1m = -2.f * sh * sh, a2 = s * frcp(foc), a3 = -a2 * foc * foc;
        const float xt = x - d;
        x = xt + (a1m * xt + a2 * px * igz) + d;
        px = px + (a3 * xt * gz + a1m * px);
    } else {
        const float foc = fsqrt(-q * igz), omg = foc * delz;
        const float sh = sinh32(0.5f * omg);
        const float a1m = 2.f * sh * sh, a2 = sinh32(omg) * frcp(foc), a3 = a2 * foc * foc;
        const float xt = x - d;
        x = xt + (a1m * xt + a2 * px * igz) + d;
        px = px + (a3 * xt * gz + a1m * px);
    }
}

__device__ __forceinline__ void transverse32(const TransParams &p, float gam, float &x, float &y,
                                             float &px, float &py) {
    if (p.corrector) {
        px += (float)p.ccx;
        py += (float)p.ccy;
    }
    if (p.chicane) {
        const float gz = fsqrt(gam * gam - 1.f - px * px - py * py), igz = frcp(gz);
        float t = x;
        x = (float)p.m00 * t + (float)p.m01 * px * igz;
        px = (float)p.m10 * t * gz + (float)p.m11 * px;
        t = y;
        y = (float)p.m22 * t + (float)p.m23 * py * igz;
        py = (float)p.m32 * t * gz + (float)p.m33 * py;
    }
    const float aw = (float)p.aw;
    const float gz = fsqrt(gam * gam - 1.f - aw * aw - px * px - py * py);
    line_map32(p.kindx, (float)p.delz, (float)p.qx, x, px, gz, (float)p.xoff);
    line_map32(p.kindy, (float)p.delz, (float)p.qy, y, py, gz, (float)p.yoff);
}

// ---------------------------------------------------------------------------
// front: first transverse half step + x/y moments (Beam::getSize,
// src/Core/Beam.cpp:42-67) + the short-range solver's centroid and largest
// radius (EFieldSolver::analyseBeam, EFieldSolver.cpp:137-157).  One
// workgroup per slice.
// ---------------------------------------------------------------------------
// Block sum of K doubles: butterfly within each wave, then the per-wave
// partials in wave order (deterministic).  lds holds K * NT/64 doubles.
template <int NT, int K>
__device__ __forceinline__ void bsum_w(double (&v)[K], double *lds) {
    constexpr int NW = NT / 64;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < K; k++)
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) lds[k * NW + wv] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        double acc = 0;
        for (int q = 0; q < NW; q++) acc += lds[k * NW + q];
        v[k] = acc;
    }
    __syncthreads();
}

// Block minimum of K ints, same structure as bsum_w; lds holds K * NT/64 ints.
template <int NT, int K>
__device__ __forceinline__ void bmin_w(int (&v)[K], int *lds) {
    constexpr int NW = NT / 64;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < K; k++)
        for (int o = 32; o > 0; o >>= 1) v[k] = min(v[k], __shfl_xor(v[k], o));
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) lds[k * NW + wv] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        int acc = lds[k * NW];
        for (int q = 1; q < NW; q++) acc = min(acc, lds[k * NW + q]);
        v[k] = acc;
    }
    __syncthreads();
}

// Block maximum of one double, same structure; lds holds NT/64 doubles.
template <int NT>
__device__ __forceinline__ double bmax_w(double v, double *lds) {
    constexpr int NW = NT / 64;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    __syncthreads();
    if (lane == 0) lds[wv] = v;
    __syncthreads();
    double acc = lds[0];
    for (int q = 1; q < NW; q++) acc = fmax(acc, lds[q]);
    __syncthreads();
    return acc;
}

constexpr int kFrontNT = 256;
// loads in flight per lane in the short particle and slice passes
constexpr int kPassU = 8;
template <int NT>
__device__ void front_finish(const MixArrays &a, int is, int moments, int sc, double *lds, double (&v)[4], int bx0,
                             int by0, int bx1, int by1, double *fsize_out);
// front of slice `is` for step S on a workgroup of NT threads (a multiple of
// 64); the slice size goes to fsize_out[is]; lds: 4 NT / 64 doubles of LDS
// scratch
template <int NT>
__device__ void front_body(const MixArrays &a, int is, int moments, int sc, double *lds, const MixStep &S,
                           double *fsize_out) {
    const int t = threadIdx.x;
    const long long b0 = a.off[is], b1 = a.off[is + 1];
    const float gref = (float)a.gref;
    double v[4] = {0, 0, 0, 0};
    const float gridmax = (float)S.p.gridmax, inv_dgrid = 1.f / (float)S.p.dgrid;
    const int ng = a.ng;
    int bx0 = 1 << 30, by0 = 1 << 30, bx1 = -1, by1 = -1;
    const TransParams t1 = S.t1;  // in registers: the stores below may alias the step table
    // kPassU particles per lane and pass, all their loads before the first
    // store: the body is short, so fewer loads in flight would leave the
    // pass waiting on HBM latency (the sums keep the particle order)
    for (long long i0 = b0 + t; i0 < b1; i0 += kPassU * NT) {
        float x4[kPassU], y4[kPassU], px4[kPassU], py4[kPassU], g4[kPassU];
#pragma unroll
        for (int u = 0; u < kPassU; u++) {
            const long long i = i0 + (long long)u * NT;
            const bool in = i < b1;
            x4[u] = in ? a.x[i] : 0.f;
            y4[u] = in ? a.y[i] : 0.f;
            px4[u] = in ? a.px[i] : 0.f;
            py4[u] = in ? a.py[i] : 0.f;
            g4[u] = in ? a.dg[i] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < kPassU; u++) {
            const long long i = i0 + (long long)u * NT;
            if (i >= b1) break;
            float x = x4[u], y = y4[u], px = px4[u], py = py4[u];
            const float gam = gref + g4[u];
            transverse32(t1, gam, x, y, px, py);
            a.x[i] = x;
            a.y[i] = y;
            a.px[i] = px;
            a.py[i] = py;
            v[0] += x;
            v[1] += (double)x * x;
            v[2] += y;
            v[3] += (double)y * y;
            if ((x > -gridmax) && (x < gridmax) && (y > -gridmax) && (y < gridmax)) {
                const int ix = min((int)floorf((x + gridmax) * inv_dgrid), ng - 2);
                const int iy = min((int)floorf((y + gridmax) * inv_dgrid), ng - 2);
                bx0 = min(bx0, ix);
                bx1 = max(bx1, ix + 1);
                by0 = min(by0, iy);
                by1 = max(by1, iy + 1);
            }
        }
    }
    front_finish<NT>(a, is, moments, sc, lds, v, bx0, by0, bx1, by1, fsize_out);
}

// The rest of a front once its particles are pushed and stored: footprint,
// slice size and centroid from the per-lane sums v (x, x^2, y, y^2) and
// cells, then the largest radius (each lane rereads the particles it wrote;
// the partition is the particle loop's, i = b0 + t + k NT).
template <int NT>
__device__ void front_finish(const MixArrays &a, int is, int moments, int sc, double *lds, double (&v)[4], int bx0,
                             int by0, int bx1, int by1, double *fsize_out) {
    const int t = threadIdx.x;
    const long long b0 = a.off[is], b1 = a.off[is + 1];
    // footprint of the slice for the main kernel's deposit tile
    int bv[4] = {bx0, by0, -bx1, -by1};
    bmin_w<NT, 4>(bv, (int *)lds);
    if (t == 0) a.pbox[is] = make_int4(bv[0], bv[1], -bv[2], -bv[3]);
    if (!moments) return;  // uniform
    bsum_w<NT, 4>(v, lds);
    const long long cnt = b1 - b0;
    double xc = 0, yc = 0;
    if (cnt > 0) {
        xc = v[0] / (double)cnt;
        yc = v[2] / (double)cnt;
    }
    if (t == 0) {
        double sz;
        if (a.cur[is] <= 0 || cnt == 0) {
            sz = 1;
        } else {
            const double c = (double)cnt;
            const double x1 = v[0] / c, x2 = v[1] / c, y1 = v[2] / c, y2 = v[3] / c;
            sz = sqrt(fabs(x2 - x1 * x1)) * sqrt(fabs(y2 - y1 * y1));
        }
        if (sz <= 0) sz = 1;
        fsize_out[is] = sz;
        a.cen[2 * is] = xc;
        a.cen[2 * is + 1] = yc;
    }
    if (!sc) return;  // uniform
    // largest radius from the centroid; each lane rereads what it wrote,
    // kPassU particles per pass with the loads first (a max, so the order is free)
    double m = 0;
    for (long long i0 = b0 + t; i0 < b1; i0 += kPassU * NT) {
        float x4[kPassU], y4[kPassU];
#pragma unroll
        for (int u = 0; u < kPassU; u++) {
            const long long i = i0 + (long long)u * NT;
            const bool in = i < b1;
            x4[u] = in ? a.x[i] : (float)xc;
            y4[u] = in ? a.y[i] : (float)yc;
        }
#pragma unroll
        for (int u = 0; u < kPassU; u++) {
            if (i0 + (long long)u * NT >= b1) break;
            const double tx = (double)x4[u] - xc, ty = (double)y4[u] - yc;
            m = fmax(m, tx * tx + ty * ty);
        }
    }
    m = bmax_w<NT>(m, lds);
    if (t == 0) a.rbound[is] = sqrt(m);
}

// the front of step istep unless the previous step's main kernel ran it
// (uniform per workgroup)
template <int NT>
__device__ __forceinline__ void front_step(const MixArrays &a, int is, int moments, int sc, double *lds) {
    const MixStep &S = a.steps[a.st->istep];
    if (!S.front_done) front_body<NT>(a, is, moments, sc, lds, S, a.mom_fsize[S.fbuf]);
}

__global__ void __launch_bounds__(kFrontNT) k_mix_front(MixArrays a, int moments, int sc) {
    __shared__ double lds[4 * (kFrontNT / 64)];
    front_step<kFrontNT>(a, blockIdx.x, moments, sc, lds);
}

// the growing radial grid of the short-range solver, in slice order
// (EFieldSolver.cpp:159-165); the same staged ballot walk as the FP64 form
__device__ __forceinline__ void scan_body(const MixArrays &a, double *lds) {
    scan_staged(a.ns, a.off, a.rbound, a.rmax, a.sc_ng, a.dr, lds);
}

__global__ void __launch_bounds__(kScanNT) k_mix_scan(MixArrays a) {
    __shared__ double lds[kScanChunk];
    scan_body(a, lds);
}

// ---------------------------------------------------------------------------
// main: one workgroup per beam slice.
//   long-range space charge of the slice (EFieldSolver::longRange :103-116)
//   short-range space charge: radial histograms and the radial tridiagonal
//     per mode (EFieldSolver::shortRange :197-258)
//   per particle: field gather (BeamSolver.cpp:66-79), RK4 (:89-141) in
//     offset form, R56 (TrackBeam::applyR56), second transverse half step,
//     beam sums (DiagBeam::getValues), footprint box
//   deposit of the slice's source (FieldSolverADI::advance :13-51) into an
//     LDS tile, written to the source plane inside the box
// One dynamic LDS region is reused: histograms + mode solutions during the
// space-charge part and the particle pass, reduction scratch after it, the
// deposit tile last.
// ---------------------------------------------------------------------------
constexpr int kScMax = 256;
constexpr int kMixDyn = 8704;    // float2 cells of the large dynamic region (68 KiB)
constexpr int kMixDynS = 5632;   // the small one (44 KiB)
constexpr int kMixDynT = 4096;   // the smallest (32 KiB)

struct OdeMix {
    float xks, eps_ref, bref, gref, ztemp1, ez;
    float2 rpart;
    float btpar_m1;  // btpar - bref = px^2 + py^2 + aw^2 (awloc^2 - 1)
    float kgg, kpp;
    __device__ __forceinline__ void eval(float dgam, float phase) {
        float s, c;
        __sincosf(phase, &s, &c);
        // rpart * (cos, -sin)
        const float cre = rpart.x * c + rpart.y * s, cim = rpart.y * c - rpart.x * s;
        const float gam = gref + dgam;
        const float ig = frcp(gam), ig2 = ig * ig;
        const float dbt = btpar_m1 + ztemp1 * cre;   // btper0 - bref
        const float eps = (bref + dbt) * ig2;
        const float ibeta = frsq(1.f - eps);          // 1 / beta
        // eps_ref - eps = bref (gam^2 - gref^2) / (gref^2 gam^2) - (btper0 - bref) / gam^2
        const float deps = (eps_ref * dgam * (2.f * gref + dgam) - dbt) * ig2;
        kpp += xks * (0.5f + 0.375f * (eps + eps_ref)) * deps;
        kgg += cim * ibeta * ig - ez;
    }
};

// a + b rounded stochastically to one of the two neighbouring FP32 values,
// with probability proportional to the distance (unbiased: E[result] = a + b
// exactly, up to the draw's resolution).  The residual of round-to-nearest
// comes from TwoSum; the top 24 bits of u are the uniform draw (the push
// passes 16 random bits and zeros).  Accumulating a small per-step increment into an
// O(1) offset with round-to-nearest repeats the same rounding error every
// step; stochastic rounding makes it a zero-mean random walk instead.
__device__ __forceinline__ float add_sr(float a, float b, unsigned u) {
    const float s = a + b;
    const float bb = s - a;
    const float err = (a - (s - bb)) + (b - bb);
    if (err == 0.f) return s;
    const float nb = __int_as_float(__float_as_int(s) + ((err > 0.f) == (s > 0.f) ? 1 : -1));
    const float gap = fabsf(nb - s);
    return ((float)(u >> 8) * 5.9604645e-08f) * gap < fabsf(err) ? nb : s;
}

// 32-bit mix of a particle index and a step counter (xorshift-multiply)
__device__ __forceinline__ unsigned mix_hash(unsigned long long i, unsigned step, unsigned salt) {
    unsigned x = (unsigned)i * 0x9E3779B9u ^ (unsigned)(i >> 32) * 0x85EBCA6Bu ^ step * 0xC2B2AE35u ^ salt;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

// radial cell of a particle: rdr = radius / cell width
__device__ __forceinline__ int sc_bin(float rdr, int ngr) {
    int bin = (int)floorf(rdr);
    return (bin >= 0 && bin < ngr) ? bin : (bin >= ngr ? ngr - 1 : 0);
}

__device__ __forceinline__ float2 cw_pow(float2 cw, int m) {
    // cw^m for a unit complex cw: conj(cw)^|m| for m < 0
    float2 p = f2(1, 0);
    const int um = m < 0 ? -m : m;
    const float2 b = m < 0 ? f2(cw.x, -cw.y) : cw;
    for (int q = 0; q < um; q++) p = fmul(p, b);
    return p;
}

// LDS accumulation is FP64 (ds_add_f64): on gfx950 FP32 LDS atomics
// (ds_add_f32) issue far slower than FP64 ones, and the sums gain precision.
// largest space-charge mode table (float2 cells); a run takes (2 nphi + 1)
// nz modes x radial cells of the main kernel's region, the deposit tile the rest
constexpr int kCelmCells = 1024;
// accumulation type of the main kernel's deposit tile (FP64: an FP32 tile
// holds twice the footprint but its atomics made the kernel 2.1x slower)
using TileT = double;

// EG: also sum the push's energy increments per slice (egsum, the energy
// balance of the learned field step).  A variant of its own because the
// extra accumulator costs registers the common path does not pay for.
// NT: threads per slice (4, 6 or 8 wavefronts); DYN: float2 cells of the
// dynamic LDS region (68 or 44 KiB: the larger deposit tile, or more
// workgroups per CU).  The autotuner picks both per context.
template <bool EG, int NT, int DYN>
__device__ bool main_step(const MixArrays &a, const int is, int sc_on, float2 *dyn) {
    // the region: the space-charge mode solutions ((2 nphi + 1) nz modes x
    // radial cells when the short-range solver runs, at most kCelmCells,
    // rounded up to a double2 boundary), then the histograms and later the
    // deposit tile in the rest
    static_assert((NT / 64) * kScMax <= (DYN - kCelmCells) / 2, "the histograms must fit in the tile region");
    const int ncelm = sc_on ? (((2 * a.sc_nphi + 1) * a.sc_nz * a.sc_ng + 1) & ~1) : 0;
    const int tile_cells = (DYN - ncelm) / 2;             // double2 cells
    // deposit tile cells: two planes of TileT
    const int dep_cells = (DYN - ncelm) * (int)sizeof(float2) / (2 * (int)sizeof(TileT));
    __shared__ int bred[5 * (NT / 64)];
    __shared__ double wred[9 * (NT / 64)];
    __shared__ float gsol[kScMax], scupp[kScMax], sclow[kScMax], scmid[kScMax];
    __shared__ float2 scrhs[kScMax];
    const MixStep &S = a.steps[a.st->istep];
    const int t = threadIdx.x, w = t >> 6;
    const long long b0 = a.off[is], b1 = a.off[is + 1];
    const int npart = (int)(b1 - b0);
    const int ns = a.ns, ng = a.ng;
    const int first = a.st->first;
    const float2 *field = a.field[a.st->cur] + (long long)((is + first) % ns) * ng * ng;
    float2 *celm = dyn;                                   // [nmodes][ngr]
    double2 *work = (double2 *)(dyn + ncelm);             // histograms / deposit tile

    // ---- long-range space charge of this slice
    double esc = 0;
    if (S.longrange) {
        double e = 0;
        const double *fsz = a.mom_fsize[S.fbuf];
        // kPassU slices per lane and pass, loads first (the same order of sums)
        for (int j0 = t; j0 < ns; j0 += kPassU * NT) {
            double sz4[kPassU], cu4[kPassU];
#pragma unroll
            for (int u = 0; u < kPassU; u++) {
                const int j = j0 + u * NT;
                const bool in = j < ns;
                sz4[u] = in ? fsz[j] : 1.0;
                cu4[u] = in ? a.cur[j] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kPassU; u++) {
                const int j = j0 + u * NT;
                if (j >= ns) break;
                const double sz = sz4[u];
                const double ds = (double)(j - is) * S.lsc_ds;
                const double coef = 1 - sqrt(ds * ds / (ds * ds + sz));
                const double sgn = (j > is) ? -1. : ((j < is) ? 1. : 0.);
                e += sgn * coef * S.lsc_scl * cu4[u] / sz;
            }
        }
        double v[1] = {e};
        bsum_w<NT, 1>(v, wred);
        esc = v[0];
    }
    if (t == 0) a.esc[is] = esc;  // the slice's LSC field for the diagnostics
    const float ez_long = (float)(-esc / kEev);

    // ---- short-range space charge
    const bool sc = sc_on && S.use_sc && npart > 0;
    const double basep = a.base[is];
    const PushParams &P = S.p;
    const float ph0 = wrap_phase(basep);
    const float phh = wrap_phase(basep + S.kappa * 0.5 * P.delz);
    const float ph1 = wrap_phase(basep + S.kappa * P.delz);
    const int ngr = a.sc_ng;
    float xc = 0, yc = 0, idr = 1;  // idr: 1 / radial cell width
    if (sc) {
        xc = (float)a.cen[2 * is];
        yc = (float)a.cen[2 * is + 1];
        idr = (float)(1.0 / a.dr[is]);
        const double econst = kVacimp / kEev * a.cur[is] / S.sc_ks / (double)npart;
        const double coef = -S.sc_gz2 / S.sc_ks / S.sc_ks;
        const double drd = a.dr[is];
        // sub-histograms per lane group (lane mod nsub) within each wavefront:
        // a beam's particles crowd into the inner radial cells, and lanes
        // adding into one cell serialise in the LDS
        const int nsub = (NT / 64) * 4 * ngr <= tile_cells ? 4 : 1;
        double2 *hs = work;  // [NT/64][nsub][ngr]
        double2 *myh = hs + (w * nsub + (t & (nsub - 1))) * ngr;
        int mode = 0;
        for (int m = -a.sc_nphi; m <= a.sc_nphi; m++) {
            for (int l = 1; l <= a.sc_nz; l++, mode++) {
                __syncthreads();
                for (int k = t; k < (NT / 64) * nsub * ngr; k += NT) hs[k] = make_double2(0, 0);
                __syncthreads();
                // kPassU particles per lane and pass, loads first: the body is
                // short, so one particle in flight per lane would leave the
                // pass waiting on HBM latency
                for (long long i0 = b0 + t; i0 < b1; i0 += kPassU * NT) {
                    float px4[kPassU], py4[kPassU], pt4[kPassU];
#pragma unroll
                    for (int u = 0; u < kPassU; u++) {
                        const long long i = i0 + (long long)u * NT;
                        const bool in = i < b1;
                        px4[u] = in ? a.x[i] : 0.f;
                        py4[u] = in ? a.y[i] : 0.f;
                        pt4[u] = in ? a.dth[i] : 0.f;
                    }
#pragma unroll
                    for (int u = 0; u < kPassU; u++) {
                        if (i0 + (long long)u * NT >= b1) break;
                        const float tx = px4[u] - xc, ty = py4[u] - yc, th = pt4[u];
                        const float radi = fsqrt(tx * tx + ty * ty), ir = frcp(radi);
                        const int bin = sc_bin(radi * idr, ngr);
                        float s, c;
                        __sincosf((float)l * (ph0 + th), &s, &c);
                        const float2 v = fmul(cw_pow(f2(tx * ir, ty * ir), -m), f2(c, -s));
                        unsafeAtomicAdd(&myh[bin].x, (double)v.x);
                        unsafeAtomicAdd(&myh[bin].y, (double)v.y);
                    }
                }
                __syncthreads();
                // constructLaplaceOperator (:176-190): coefficients and source of
                // every radial cell, one cell per lane; tridiag (:260-272) on
                // lane 0.  The operator is real, the right-hand side complex.
                for (int i = t; i < ngr; i += NT) {
                    const double voli = kPi * drd * drd * (i == 0 ? 1.0 : (double)(2 * i + 1));
                    const double ldi = i > 0 ? 2 * kPi * i : 0.0;
                    const double ldn = (i + 1 < ngr) ? 2 * kPi * (i + 1) : 0.0;
                    const double rl = i == 0 ? 0.5 : log((double)(i + 1) / (double)i);
                    double lm = -ldi - ldn - 2. * kPi * m * m * rl;
                    if (i == ngr - 1) lm -= 2 * kPi * ngr;
                    scmid[i] = (float)(1 + coef * lm / l / l / voli);
                    if (i > 0) {
                        const double volm = kPi * drd * drd * (i == 1 ? 1.0 : (double)(2 * i - 1));
                        scupp[i] = (float)(coef * ldi / l / l / volm);
                        sclow[i] = (float)(coef * ldi / l / l / voli);
                    }
                    double csx = 0, csy = 0;
                    for (int q = 0; q < (NT / 64) * nsub; q++) {
                        csx += hs[q * ngr + i].x;
                        csy += hs[q * ngr + i].y;
                    }
                    const double e = econst / l / voli;
                    scrhs[i] = f2((float)(-csy * e), (float)(csx * e));
                }
                __syncthreads();
                if (t == 0) {
                    float2 *ce = celm + mode * ngr;
                    float bet = scmid[0];
                    ce[0] = fscale(scrhs[0], 1.f / bet);
                    for (int i = 1; i < ngr; i++) {
                        gsol[i] = scupp[i] / bet;
                        bet = scmid[i] - sclow[i] * gsol[i];
                        ce[i] = fscale(fsub(scrhs[i], fscale(ce[i - 1], sclow[i])), 1.f / bet);
                    }
                    for (int i = ngr - 2; i > -1; i--) ce[i] = fsub(ce[i], fscale(ce[i + 1], gsol[i + 1]));
                    if (l == 1) a.efield[is] = hypot((double)ce[0].x, (double)ce[0].y);
                }
            }
        }
    }

    // ---- deposit tile: the slice's footprint after the first transverse half
    // step (front kernel) padded by 2 cells.  The second half step moves a
    // particle by |px/gamma_z| dz/2, far below a cell; a particle that does
    // leave the tile sets the overflow flag and the slice then takes the
    // exact two-pass deposit below.
    const DepositParams &D = S.d;
    const bool dep = D.enabled != 0;
    const int4 pb = a.pbox[is];
    int tx0 = 0, ty0 = 0, tw = 0, th = 0;
    bool fused = false;
    if (dep && pb.z >= pb.x) {
        tx0 = max(0, pb.x - 2);
        ty0 = max(0, pb.y - 2);
        tw = min(ng - 1, pb.z + 2) - tx0 + 1;
        th = min(ng - 1, pb.w + 2) - ty0 + 1;
        fused = tw * th <= dep_cells;
    }
    float scl = 0;
    if (dep) {
        const double scl0 = D.fcoup * kVacimp * a.cur[is] * D.xks * D.delz;
        scl = npart > 0 ? (float)(scl0 / (4 * kEev * (double)npart * D.dgrid * D.dgrid)) : 0.f;
    }
    __syncthreads();  // histograms dead, tile takes their place
    // the tile as two planes (real parts, then imaginary parts): a lane's
    // atomic then spans fewer LDS banks, and the random cells of a
    // wavefront's particles collide in fewer banks
    TileT *tre = (TileT *)work, *tim = tre + tw * th;
    if (fused)
        for (int k = t; k < 2 * tw * th; k += NT) tre[k] = (TileT)0;
    __syncthreads();

    // ---- particle pass
    const float gref = (float)a.gref;
    const float xks = (float)P.xks;
    const float bref = (float)(1 + P.aw * P.aw);
    const float aw2 = (float)(P.aw * P.aw);
    const float gridmax = (float)P.gridmax, dgrid = (float)P.dgrid;
    const float rtmp = (float)P.rtmp;
    const float delz = (float)P.delz;
    const float dgam0 = (float)(a.gref - S.gamma0);  // gamma - gamma0 = dg + dgam0
    const float r56 = (float)S.r56;
    const float autoph = (float)P.autophase;
    const float dax = (float)D.ax, day = (float)D.ay, dkx = (float)D.kx, dky = (float)D.ky;
    const float dgx = (float)D.gradx, dgy = (float)D.grady;
    // step constants in registers: the particle stores below may alias the
    // step table as far as the compiler can tell, which would reload them
    const float pax = (float)P.ax, pay = (float)P.ay, pkx = (float)P.kx, pky = (float)P.ky;
    const float pgx = (float)P.gradx, pgy = (float)P.grady;
    const float eps_ref = (float)S.eps_ref;
    const bool chic = S.chicane != 0;
    const TransParams t2 = S.t2;
    // With fnext and the fused tile the next step's front rides in this loop:
    // after the deposit, the first transverse half step of step k+1 on the
    // registers, its slice sums and footprint, one store.  The post-t2
    // positions also go to xs/ys for the two-pass deposit, should a particle
    // leave the tile.  (Without the fused tile the front runs after the
    // deposit, as a second pass.)
    const bool nextf = S.fnext && fused;
    const TransParams t1n = a.steps[a.st->istep + (nextf ? 1 : 0)].t1;
    double nv[4] = {0, 0, 0, 0};
    int nbx0 = 1 << 30, nby0 = 1 << 30, nbx1 = -1, nby1 = -1;
    // per-lane beam sums of at most a few dozen particles in FP32, energy as
    // the offset dg; widened to FP64 for the block sum
    constexpr int NV = EG ? 9 : 8;
    float fv[NV];  // beam sums; [8]: the push's energy increments
#pragma unroll
    for (int k = 0; k < NV; k++) fv[k] = 0.f;
    int bx0 = 1 << 30, by0 = 1 << 30, bx1 = -1, by1 = -1, ovf = 0;
    const unsigned stepno = a.st->seq;
    const float inv_dgrid = 1.f / dgrid;
    // the next particle's coordinates are loaded before this one's stores,
    // so their latency overlaps the RK4 of the current particle
    float n_th = 0, n_g = 0, n_x = 0, n_y = 0, n_px = 0, n_py = 0;
    if (b0 + t < b1) {
        const long long i = b0 + t;
        n_th = a.dth[i];
        n_g = a.dg[i];
        n_x = a.x[i];
        n_y = a.y[i];
        n_px = a.px[i];
        n_py = a.py[i];
    }
    for (long long i = b0 + t; i < b1; i += NT) {
        const float th_in = n_th;
        const float g0 = n_g, t0 = th_in + autoph;
        float x = n_x, y = n_y, px = n_px, py = n_py;
        if (i + NT < b1) {
            const long long j = i + NT;
            n_th = a.dth[j];
            n_g = a.dg[j];
            n_x = a.x[j];
            n_y = a.y[j];
            n_px = a.px[j];
            n_py = a.py[j];
        }
        const float dx = x - pax, dy = y - pay;
        const float awm1 = 0.5f * (pkx * dx * dx + pky * dy * dy) + pgx * dx + pgy * dy;  // awloc - 1
        OdeMix o;
        o.xks = xks;
        o.eps_ref = eps_ref;
        o.bref = bref;
        o.gref = gref;
        o.ztemp1 = -2.f / xks;
        o.btpar_m1 = px * px + py * py + aw2 * awm1 * (awm1 + 2.f);
        float ez = ez_long;
        if (sc) {
            const float tx = x - xc, ty = y - yc;
            const float radi = fsqrt(tx * tx + ty * ty), ir = frcp(radi);
            const float2 cw = f2(tx * ir, ty * ir);
            const int bin = sc_bin(radi * idr, ngr);
            int mode = 0;
            for (int m = -a.sc_nphi; m <= a.sc_nphi; m++) {
                for (int l = 1; l <= a.sc_nz; l++, mode++) {
                    float s, c;
                    __sincosf((float)l * (ph0 + th_in), &s, &c);
                    const float2 ct = fmul(cw_pow(cw, m), f2(c, s));
                    ez += 2.f * fmul(ct, celm[mode * ngr + bin]).x;
                }
            }
        }
        o.ez = ez;
        o.rpart = f2(0, 0);
        if ((x > -gridmax) && (x < gridmax) && (y > -gridmax) && (y < gridmax)) {
            float wx = (x + gridmax) * inv_dgrid, wy = (y + gridmax) * inv_dgrid;
            const int ix = min((int)floorf(wx), ng - 2), iy = min((int)floorf(wy), ng - 2);
            wx = 1.f + floorf(wx) - wx;
            wy = 1.f + floorf(wy) - wy;
            const int idx = ix + iy * ng;
            const float2 q0 = field[idx], q1 = field[idx + 1], q2 = field[idx + ng],
                         q3 = field[idx + ng + 1];
            float2 cp = fscale(q0, wx * wy);
            cp = fadd(cp, fscale(q1, (1 - wx) * wy));
            cp = fadd(cp, fscale(q2, wx * (1 - wy)));
            cp = fadd(cp, fscale(q3, (1 - wx) * (1 - wy)));
            const float r = rtmp * (1.f + awm1);
            o.rpart = f2(r * cp.x, -r * cp.y);
        }
        // RungeKutta (BeamSolver.cpp:93-141): the reference updates the state
        // in place between evaluations; here the increments (eg, et) follow
        // that same sequence from zero, the stages are evaluated at
        // (g0 + eg, t0 + et) on the base phases base, base + kappa dz/2
        // (twice) and base + kappa dz, and the step's total increment is
        // added to the stored offsets once, with stochastic rounding.
        float k3gg, k3pp, eg = 0.f, et = 0.f;
        o.kgg = 0;
        o.kpp = 0;
        o.eval(g0, ph0 + t0);
        float stpz = 0.5f * delz;
        eg += stpz * o.kgg;
        et += stpz * o.kpp;
        k3gg = o.kgg;
        k3pp = o.kpp;
        o.kgg = 0;
        o.kpp = 0;
        o.eval(g0 + eg, phh + (t0 + et));
        eg += stpz * (o.kgg - k3gg);
        et += stpz * (o.kpp - k3pp);
        k3gg *= 1.f / 6.f;
        k3pp *= 1.f / 6.f;
        o.kgg *= -0.5f;
        o.kpp *= -0.5f;
        o.eval(g0 + eg, phh + (t0 + et));
        stpz = delz;
        eg += stpz * o.kgg;
        et += stpz * o.kpp;
        k3gg -= o.kgg;
        k3pp -= o.kpp;
        o.kgg *= 2;
        o.kpp *= 2;
        o.eval(g0 + eg, ph1 + (t0 + et));
        eg += stpz * (k3gg + o.kgg * (1.f / 6.f));
        et += stpz * (k3pp + o.kpp * (1.f / 6.f));
        // one hash for both roundings, 16 bits each (the rounding
        // probability then resolves 2^-16 of an ulp)
        const unsigned hs = mix_hash((unsigned long long)i, stepno, 0x51u);
        const float dg = add_sr(g0, eg, hs << 16);
        if constexpr (EG) fv[NV - 1] += eg;
        if (chic) et += r56 * (dg + dgam0);  // TrackBeam::applyR56
        const float dth = add_sr(t0, et, hs & 0xFFFF0000u);
        transverse32(t2, gref + dg, x, y, px, py);
        a.dg[i] = dg;
        a.dth[i] = dth;
        // beam sums of the updated state
        fv[0] += dg;
        fv[1] += dg * dg;
        fv[2] += x;
        fv[3] += x * x;
        fv[4] += y;
        fv[5] += y * y;
        float s, c;
        __sincosf(ph1 + dth, &s, &c);
        fv[6] += c;
        fv[7] += s;
        if (dep && (x > -gridmax) && (x < gridmax) && (y > -gridmax) && (y < gridmax)) {
            float wx = (x + gridmax) * inv_dgrid, wy = (y + gridmax) * inv_dgrid;
            const int ix = min((int)floorf(wx), ng - 2), iy = min((int)floorf(wy), ng - 2);
            bx0 = min(bx0, ix);
            bx1 = max(bx1, ix + 1);
            by0 = min(by0, iy);
            by1 = max(by1, iy + 1);
            if (fused) {
                if (ix >= tx0 && ix + 1 < tx0 + tw && iy >= ty0 && iy + 1 < ty0 + th) {
                    wx = 1.f + floorf(wx) - wx;
                    wy = 1.f + floorf(wy) - wy;
                    const float ddx = x - dax, ddy = y - day;
                    const float faw2 = 1.f + dkx * ddx * ddx + dky * ddy * ddy + 2.f * (dgx * ddx + dgy * ddy);
                    const float part = fsqrt(faw2) * scl * frcp(gref + dg);
                    const float2 cp = f2(s * part, c * part);
                    const int lk = (iy - ty0) * tw + (ix - tx0);
                    unsafeAtomicAdd(&tre[lk], (TileT)(wx * wy * cp.x));
                    unsafeAtomicAdd(&tim[lk], (TileT)(wx * wy * cp.y));
                    unsafeAtomicAdd(&tre[lk + 1], (TileT)((1 - wx) * wy * cp.x));
                    unsafeAtomicAdd(&tim[lk + 1], (TileT)((1 - wx) * wy * cp.y));
                    unsafeAtomicAdd(&tre[lk + tw], (TileT)(wx * (1 - wy) * cp.x));
                    unsafeAtomicAdd(&tim[lk + tw], (TileT)(wx * (1 - wy) * cp.y));
                    unsafeAtomicAdd(&tre[lk + tw + 1], (TileT)((1 - wx) * (1 - wy) * cp.x));
                    unsafeAtomicAdd(&tim[lk + tw + 1], (TileT)((1 - wx) * (1 - wy) * cp.y));
                } else {
                    ovf = 1;
                }
            }
        }
        if (nextf) {
            a.xs[i] = x;
            a.ys[i] = y;
            transverse32(t1n, gref + dg, x, y, px, py);
            nv[0] += x;
            nv[1] += (double)x * x;
            nv[2] += y;
            nv[3] += (double)y * y;
            if ((x > -gridmax) && (x < gridmax) && (y > -gridmax) && (y < gridmax)) {
                const int ix = min((int)floorf((x + gridmax) * inv_dgrid), ng - 2);
                const int iy = min((int)floorf((y + gridmax) * inv_dgrid), ng - 2);
                nbx0 = min(nbx0, ix);
                nbx1 = max(nbx1, ix + 1);
                nby0 = min(nby0, iy);
                nby1 = max(nby1, iy + 1);
            }
        }
        a.x[i] = x;
        a.y[i] = y;
        a.px[i] = px;
        a.py[i] = py;
    }
    if (nextf) {
        const MixStep &N = a.steps[a.st->istep + 1];
        front_finish<NT>(a, is, (sc_on || S.longrange) ? 1 : 0, sc_on, wred, nv, nbx0, nby0, nbx1, nby1,
                         a.mom_fsize[N.fbuf]);
    }
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = fv[k];
    bsum_w<NT, NV>(v, wred);
    if (t == 0) {
        // sum gamma = n gref + sum dg, sum gamma^2 = n gref^2 + 2 gref sum dg + sum dg^2
        const double n = (double)npart, gr = a.gref;
        a.beamsum[8 * is + 0] = n * gr + v[0];
        a.beamsum[8 * is + 1] = n * gr * gr + 2.0 * gr * v[0] + v[1];
        for (int k = 2; k < 8; k++) a.beamsum[8 * is + k] = v[k];
        if constexpr (EG) a.egsum[is] = v[NV - 1];
        a.base[is] = basep + S.kappa * P.delz;
    }
    // ---- footprint box and overflow flag
    int bv[5] = {bx0, by0, -bx1, -by1, -ovf};
    bmin_w<NT, 5>(bv, bred);
    bx0 = bv[0];
    by0 = bv[1];
    bx1 = -bv[2];
    by1 = -bv[3];
    ovf = -bv[4];
    if (!dep || bx1 < 0) {
        if (t == 0) a.box[is] = make_int4(1, 1, 0, 0);
        return nextf;
    }
    float2 *plane = a.src + (long long)is * ng * ng;
    if (fused && !ovf) {
        for (int k = t; k < tw * th; k += NT)
            plane[(ty0 + k / tw) * ng + tx0 + k % tw] = f2((float)tre[k], (float)tim[k]);
        if (t == 0) a.box[is] = make_int4(tx0, ty0, tx0 + tw - 1, ty0 + th - 1);
        return nextf;
    }
    // ---- exact two-pass deposit over the footprint box (rows in bands of the tile)
    bx1 = min(bx1, ng - 1);
    by1 = min(by1, ng - 1);
    if (t == 0) a.box[is] = make_int4(bx0, by0, bx1, by1);
    const int ww = bx1 - bx0 + 1;
    const int rows_per_band = max(1, dep_cells / ww);
    for (int r0 = by0; r0 <= by1; r0 += rows_per_band) {
        const int r1 = min(by1, r0 + rows_per_band - 1);
        const int ncell = (r1 - r0 + 1) * ww;
        TileT *pre = (TileT *)work, *pim = pre + ncell;  // two planes, as the fused tile
        for (int k = t; k < 2 * ncell; k += NT) pre[k] = (TileT)0;
        __syncthreads();
        for (long long i = b0 + t; i < b1; i += NT) {
            const float x = nextf ? a.xs[i] : a.x[i], y = nextf ? a.ys[i] : a.y[i];  // post-t2
            if (!((x > -gridmax) && (x < gridmax) && (y > -gridmax) && (y < gridmax))) continue;
            float wx = (x + gridmax) * inv_dgrid, wy = (y + gridmax) * inv_dgrid;
            const int ix = min((int)floorf(wx), ng - 2), iy = min((int)floorf(wy), ng - 2);
            if (iy + 1 < r0 || iy > r1) continue;
            wx = 1.f + floorf(wx) - wx;
            wy = 1.f + floorf(wy) - wy;
            const float ddx = x - dax, ddy = y - day;
            const float faw2 = 1.f + dkx * ddx * ddx + dky * ddy * ddy + 2.f * (dgx * ddx + dgy * ddy);
            const float part = fsqrt(faw2) * scl * frcp(gref + a.dg[i]);
            float s, c;
            __sincosf(ph1 + a.dth[i], &s, &c);
            const float2 cp = f2(s * part, c * part);
            const float wgt[4] = {wx * wy, (1 - wx) * wy, wx * (1 - wy), (1 - wx) * (1 - wy)};
            const int cx[4] = {ix, ix + 1, ix, ix + 1};
            const int cy[4] = {iy, iy, iy + 1, iy + 1};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (cy[k] < r0 || cy[k] > r1 || cx[k] < bx0 || cx[k] > bx1) continue;
                const int lk = (cy[k] - r0) * ww + (cx[k] - bx0);
                unsafeAtomicAdd(&pre[lk], (TileT)(wgt[k] * cp.x));
                unsafeAtomicAdd(&pim[lk], (TileT)(wgt[k] * cp.y));
            }
        }
        __syncthreads();
        for (int k = t; k < ncell; k += NT) {
            const int ry = r0 + k / ww, rx = bx0 + k % ww;
            plane[ry * ng + rx] = f2((float)pre[k], (float)pim[k]);
        }
        __syncthreads();
    }
    return nextf;
}

// main of slice is; with fnext, the next step's front of the slice either
// rides in the particle loop (main_step returns true) or runs after it on
// the particles this workgroup has just stored (each thread rereads its own:
// the front partitions a slice as the main loop does).  The slice size goes
// to the next step's size buffer: other workgroups of this launch still read
// this step's.
template <bool EG, int NT, int DYN>
__device__ void main_body(const MixArrays &a, const int is, int sc_on, float2 *dyn) {
    const bool done = main_step<EG, NT, DYN>(a, is, sc_on, dyn);
    const int k = a.st->istep;
    if (a.steps[k].fnext && !done) {
        __syncthreads();  // the deposit tile's last reads before its LDS is reused
        const int moments = (sc_on || a.steps[k].longrange) ? 1 : 0;
        const MixStep &N = a.steps[k + 1];
        front_body<NT>(a, is, moments, sc_on, (double *)dyn, N, a.mom_fsize[N.fbuf]);
    }
}

template <bool EG, int NT, int DYN>
__global__ void __launch_bounds__(NT) k_mix_main(MixArrays a, int sc_on) {
    extern __shared__ float2 dyn[];
    main_body<EG, NT, DYN>(a, blockIdx.x, sc_on, dyn);
}

// ---------------------------------------------------------------------------
// ADI of the mixed form: both half steps (FieldSolverADI::ADI :60-91) for a
// 32x32 output tile, with the truncated banded inverse B ~ A^-1 in FP32
// (dropped entries < 2^-28 of the diagonal).
//
// REF (FP64 residual correction): each half step solves A u = r as
//   u1 = B r (FP32),  rho = r - A u1 (FP64),  u = u1 + B rho,
// with r and u kept in FP64, so the step is accurate to ~(2^-24)^2 before
// the result is rounded to the FP32 field.  The correction needs r on a
// halo of 2M+1 cells instead of M.
// Without REF the half steps are u = B r in FP32.
// ---------------------------------------------------------------------------
template <int M, bool REF>
struct MixGeom {
    static constexpr int TS = 32;
    static constexpr int W = 2 * M + 1;
    static constexpr int H = REF ? 2 * M + 1 : M;   // halo of the E region
    static constexpr int RE = TS + 2 * H;
    static constexpr int PE = ((RE + 6) / 8) * 8 + 1;
    static constexpr int PG = TS + 1;
    static constexpr int CX = REF ? (M <= 4 ? 4 : 2) : 8;   // x-stage outputs per task
    static constexpr int RY = REF ? (M <= 4 ? 4 : 2) : 4;   // y-stage outputs per task
    static constexpr int BR = REF ? M + 1 : 0;      // extra band rows each side
    static constexpr int NB = TS + 2 * BR;          // band rows staged for edge tiles
    using ET = typename std::conditional<REF, double2, float2>::type;
    static constexpr int SR = (RE + (256 / RE) - 1) / (256 / RE);
    static constexpr size_t lds_bytes =
        (size_t)RE * PE * sizeof(ET) + (size_t)RE * PG * sizeof(ET) + 2 * (size_t)NB * W * sizeof(float2);
};

template <int M, bool REF>
__global__ void __launch_bounds__(256) k_mix_adi(MixArrays a) {
    using Gm = MixGeom<M, REF>;
    using ET = typename Gm::ET;
    constexpr int TS = Gm::TS, W = Gm::W, H = Gm::H, RE = Gm::RE, PE = Gm::PE, PG = Gm::PG;
    constexpr int CX = Gm::CX, RY = Gm::RY, BR = Gm::BR, NB = Gm::NB;
    constexpr int NT = 256;
    extern __shared__ double2 smd[];
    ET *E = (ET *)smd;                 // [RE][PE]
    ET *G = E + RE * PE;               // [RE][PG]
    float2 *BX = (float2 *)(G + RE * PG);  // [NB][W]
    float2 *BY = BX + NB * W;          // [NB][W]
    const int t = threadIdx.x;
    const MixStep &S = a.steps[a.st->istep];
    const MixBand bi = a.binfo[S.band];
    const float2 *band = a.bands + bi.off;
    const double aa = bi.a;
    const int ng = a.ng, ns = a.ns;
    const int ntx = (ng + TS - 1) / TS, ntiles = ntx * ntx;
    // XCD-aware order (see k_adi_band): all tiles of a slice on one XCD
    const int b = blockIdx.x;
    const int ii = (b / (8 * ntiles)) * 8 + (b & 7);
    const int tile = (b % (8 * ntiles)) >> 3;
    if (ii >= ns) return;
    const int tx0 = (tile % ntx) * TS, ty0 = (tile / ntx) * TS;
    const int first = a.st->first;
    const long long fs = (long long)((ii + first) % ns);
    const int cur = a.st->cur;
    const float2 *u = a.field[cur] + fs * ng * ng;
    float2 *uo = a.field[cur ^ 1] + fs * ng * ng;
    const float2 *Sp = a.src + (long long)ii * ng * ng;
    const int4 bx = a.box[ii];
    auto sat = [&](int x, int y) {
        return (x >= bx.x && x <= bx.z && y >= bx.y && y <= bx.w) ? Sp[y * ng + x] : f2(0, 0);
    };
    const bool xin = tx0 - BR >= bi.klo && tx0 + TS - 1 + BR <= bi.khi;
    const bool yin = ty0 - BR >= bi.klo && ty0 + TS - 1 + BR <= bi.khi;
    const float2 *bc = band + (long long)((bi.klo + bi.khi) / 2) * W;
    if (!xin)
        for (int e = t; e < NB * W; e += NT) {
            const int x = tx0 - BR + e / W;
            BX[e] = (x >= 0 && x < ng) ? band[(long long)x * W + e % W] : f2(0, 0);
        }
    if (!yin)
        for (int e = t; e < NB * W; e += NT) {
            const int y = ty0 - BR + e / W;
            BY[e] = (y >= 0 && y < ng) ? band[(long long)y * W + e % W] : f2(0, 0);
        }
    // ---- E = S + (1 + i a d_y^2) u on the RE x RE region (zero off the grid)
    {
        constexpr int SR = Gm::SR;
        constexpr int NSTRIP = (RE + SR - 1) / SR;
        if (t < RE * NSTRIP) {
            const int c = t % RE, r0 = (t / RE) * SR;
            const int x = tx0 - H + c;
            const bool xok = x >= 0 && x < ng;
            float2 uv[SR + 2], sv[SR];
#pragma unroll
            for (int q = 0; q < SR + 2; q++) {
                const int y = ty0 - H + r0 + q - 1;
                uv[q] = (xok && y >= 0 && y < ng && r0 + q - 1 < RE + 1) ? u[y * ng + x] : f2(0, 0);
            }
#pragma unroll
            for (int q = 0; q < SR; q++) sv[q] = (r0 + q < RE) ? sat(x, ty0 - H + r0 + q) : f2(0, 0);
#pragma unroll
            for (int q = 0; q < SR; q++) {
                const int r = r0 + q, y = ty0 - H + r;
                if (r < RE) {
                    ET v{};
                    if (xok && y >= 0 && y < ng) {
                        if constexpr (REF) {
                            const double2 uc = tod(uv[q + 1]), up = tod(uv[q + 2]), dn = tod(uv[q]);
                            const double sx = up.x - 2.0 * uc.x + dn.x, sy = up.y - 2.0 * uc.y + dn.y;
                            v = d2((double)sv[q].x + uc.x - aa * sy, (double)sv[q].y + uc.y + aa * sx);
                        } else {
                            const float af = (float)aa;
                            const float2 uc = uv[q + 1];
                            const float sx = uv[q + 2].x - 2.f * uc.x + uv[q].x;
                            const float sy = uv[q + 2].y - 2.f * uc.y + uv[q].y;
                            v = f2(sv[q].x + uc.x - af * sy, sv[q].y + uc.y + af * sx);
                        }
                    }
                    E[r * PE + c] = v;
                }
            }
        }
    }
    __syncthreads();
    // ---- implicit x on every row of the region, output columns of the tile;
    // then F = S + (1 + i a d_x^2) u' = S + 2 u' - E into G.  The band
    // accessor is chosen once per tile: interior tiles read the shared row
    // (uniform, scalar loads), edge tiles the rows staged in LDS.
    auto xstage = [&](auto bx_row) {
    for (int task = t; task < RE * (TS / CX); task += NT) {
        const int r = task % RE, g = task / RE;
        const int y = ty0 - H + r;
        const int x0 = tx0 + g * CX;
        if (y < 0 || y >= ng) {
#pragma unroll
            for (int j = 0; j < CX; j++) G[r * PG + g * CX + j] = ET{};
            continue;
        }
        if constexpr (!REF) {
            float2 ev[CX + 2 * M];
#pragma unroll
            for (int k = 0; k < CX + 2 * M; k++) ev[k] = E[r * PE + g * CX + k];
#pragma unroll
            for (int j = 0; j < CX; j++) {
                const int x = x0 + j;
                float2 acc = f2(0, 0);
#pragma unroll
                for (int m = 0; m < W; m++) acc = ffma(bx_row(x, m), ev[j + m], acc);
                float2 f = f2(0, 0);
                if (x < ng) {
                    const float2 sv = sat(x, y);
                    f = f2(sv.x + 2.f * acc.x - ev[j + M].x, sv.y + 2.f * acc.y - ev[j + M].y);
                }
                G[r * PG + g * CX + j] = f;
            }
        } else {
            constexpr int NE = CX + 4 * M + 2, NU = CX + 2 * M + 2, NR = CX + 2 * M;
            double2 ev[NE];
#pragma unroll
            for (int k = 0; k < NE; k++) ev[k] = E[r * PE + g * CX + k];
            float2 u1[NU];
#pragma unroll
            for (int k = 0; k < NU; k++) {
                const int xp = x0 - (M + 1) + k;
                float2 acc = f2(0, 0);
                if (xp >= 0 && xp < ng) {
#pragma unroll
                    for (int m = 0; m < W; m++) acc = ffma(bx_row(xp, m), tof(ev[k + m]), acc);
                }
                u1[k] = acc;
            }
            float2 rho[NR];
#pragma unroll
            for (int k = 0; k < NR; k++) {
                // rho = r - A u1, A u = (1 + 2 i a) u - i a (u_{+1} + u_{-1})
                const double2 c0 = tod(u1[k + 1]), cl = tod(u1[k]), cr = tod(u1[k + 2]);
                const double nbx = cl.x + cr.x, nby = cl.y + cr.y;
                const double aux = c0.x - 2.0 * aa * c0.y + aa * nby;
                const double auy = c0.y + 2.0 * aa * c0.x - aa * nbx;
                rho[k] = tof(d2(ev[k + M + 1].x - aux, ev[k + M + 1].y - auy));
            }
#pragma unroll
            for (int j = 0; j < CX; j++) {
                const int x = x0 + j;
                float2 d = f2(0, 0);
#pragma unroll
                for (int m = 0; m < W; m++) d = ffma(bx_row(x, m), rho[j + m], d);
                double2 f = d2(0, 0);
                if (x < ng) {
                    const float2 sv = sat(x, y);
                    const double upx = (double)u1[j + M + 1].x + (double)d.x;
                    const double upy = (double)u1[j + M + 1].y + (double)d.y;
                    f = d2((double)sv.x + 2.0 * upx - ev[j + H].x, (double)sv.y + 2.0 * upy - ev[j + H].y);
                }
                G[r * PG + g * CX + j] = f;
            }
        }
    }
    };
    if (xin) xstage([&](int, int m) -> float2 { return bc[m]; });
    else xstage([&](int x, int m) -> float2 { return BX[(x - (tx0 - BR)) * W + m]; });
    __syncthreads();
    // ---- implicit y, output rows of the tile; field sums in FP64
    const double shift = -0.5 * (double)(ng - 1);
    const int icen = (ng / 2) * ng + (ng / 2);
    double dg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    auto ystage = [&](auto by_row) {
    for (int task = t; task < TS * (TS / RY); task += NT) {
        const int c = task % TS, q = task / TS;
        const int x = tx0 + c, y0 = ty0 + q * RY;
        float2 out[RY];
        if constexpr (!REF) {
            float2 fv[RY + 2 * M];
#pragma unroll
            for (int k = 0; k < RY + 2 * M; k++) fv[k] = G[(q * RY + k) * PG + c];
#pragma unroll
            for (int j = 0; j < RY; j++) {
                float2 acc = f2(0, 0);
#pragma unroll
                for (int m = 0; m < W; m++) acc = ffma(by_row(y0 + j, m), fv[j + m], acc);
                out[j] = acc;
            }
        } else {
            constexpr int NE = RY + 4 * M + 2, NU = RY + 2 * M + 2, NR = RY + 2 * M;
            double2 gv[NE];
#pragma unroll
            for (int k = 0; k < NE; k++) gv[k] = G[(q * RY + k) * PG + c];
            float2 u1[NU];
#pragma unroll
            for (int k = 0; k < NU; k++) {
                const int yp = y0 - (M + 1) + k;
                float2 acc = f2(0, 0);
                if (yp >= 0 && yp < ng) {
#pragma unroll
                    for (int m = 0; m < W; m++) acc = ffma(by_row(yp, m), tof(gv[k + m]), acc);
                }
                u1[k] = acc;
            }
            float2 rho[NR];
#pragma unroll
            for (int k = 0; k < NR; k++) {
                const double2 c0 = tod(u1[k + 1]), cl = tod(u1[k]), cr = tod(u1[k + 2]);
                const double nbx = cl.x + cr.x, nby = cl.y + cr.y;
                const double aux = c0.x - 2.0 * aa * c0.y + aa * nby;
                const double auy = c0.y + 2.0 * aa * c0.x - aa * nbx;
                rho[k] = tof(d2(gv[k + M + 1].x - aux, gv[k + M + 1].y - auy));
            }
#pragma unroll
            for (int j = 0; j < RY; j++) {
                float2 d = f2(0, 0);
#pragma unroll
                for (int m = 0; m < W; m++) d = ffma(by_row(y0 + j, m), rho[j + m], d);
                out[j] = tof(d2((double)u1[j + M + 1].x + (double)d.x, (double)u1[j + M + 1].y + (double)d.y));
            }
        }
#pragma unroll
        for (int j = 0; j < RY; j++) {
            const int y = y0 + j;
            if (x < ng && y < ng) {
                const float2 o = out[j];
                uo[y * ng + x] = o;
                const double ox = o.x, oy = o.y;
                const double wei = ox * ox + oy * oy;
                const double dx = (double)x + shift, dy = (double)y + shift;
                dg[0] += wei;
                dg[1] += dx * wei;
                dg[2] += dx * dx * wei;
                dg[3] += dy * wei;
                dg[4] += dy * dy * wei;
                dg[5] += ox;
                dg[6] += oy;
                if (y * ng + x == icen) dg[7] = wei;
            }
        }
    }
    };
    if (yin) ystage([&](int, int m) -> float2 { return bc[m]; });
    else ystage([&](int y, int m) -> float2 { return BY[(y - (ty0 - BR)) * W + m]; });
    double *red = (double *)smd;
    bsum<NT, 8>(dg, red);
    if (t == 0) {
        double *o = a.tilediag + (fs * ntiles + tile) * 8;
        for (int k = 0; k < 8; k++) o[k] = dg[k];
    }
}

// ---------------------------------------------------------------------------
// Streaming ADI of the mixed form: both half steps of FieldSolverADI::ADI
// (:60-91) with the FP32 banded inverse B ~ A^-1, organised as one wavefront
// per strip of columns walking down the rows of a slice.  Lane l owns column
// x = x0 - H + l; lanes [H, 64-H) produce output.  For each row r:
//   E(r) = S + u + i a (u(r+1) - 2 u(r) + u(r-1))     (field rows prefetched)
//   u'(r) = sum_m B[x][m] E(r, x+m-M)                   (E row through LDS)
//   F(r) = S + 2 u'(r) - E(r)                            (kept in a register ring)
//   u''(r-M) = sum_m B[r-M][m] F(r-2M+m)                 (ring of 2M+1 rows)
// Rows and columns whose band equals the central one use it in the
// symmetric form b_M F_k + sum_j b_{M+j} (F_{k-j} + F_{k+j}) (A, and so its
// Toeplitz interior inverse, is symmetric), M + 1 complex multiplies
// instead of 2M + 1.
//
// REF: one FP64 residual-correction sweep per half step.  A = tridiag(-i a,
// 1 + 2 i a, -i a) (FieldSolverADI::init, :137-165) is diagonally dominant,
// |off-diagonal / diagonal| = a / |1 + 2 i a| ~ 0.01, so one Jacobi step on
// the FP64 residual,
//   u = u1 + D^-1 (r - A u1),  u1 = B r (FP32),  D = 1 + 2 i a,
// with r, A u1 and the update in FP64, shrinks the solve's error (band
// truncation and FP32 accumulation) by 2 a / |1 + 2 i a| before the result
// is rounded to the FP32 field.  E is exact in FP64 in its lane; F is kept as
// an FP32 value plus its FP32 remainder (F = hi + lo to ~2^-48), so the y
// residual uses F to FP64 accuracy.  The x residual needs u1 of the
// neighbouring columns (H = M + 1 halo lanes), the y residual u'' of the row
// below, so the corrected row y leaves the loop one row later.
//
// Each field value is read once (plus the strip's 2H-column halo) and
// written once.  The x band rows of edge strips are per lane (registers);
// the y band rows are uniform per row (scalar loads).
// ---------------------------------------------------------------------------
template <int M, bool REF>
struct StreamGeom {
    static constexpr int W = 2 * M + 1;
    static constexpr int H = REF ? M + 1 : M;  // halo lanes on each side
    static constexpr int SO = 64 - 2 * H;      // output columns per strip
    // FIX (narrow bands): the row loop is unrolled by W; the F ring and its
    // remainders have W slots and the field queues QN slots (QN divides W),
    // so every slot of row r is the fixed register r mod W (r mod QN) and no
    // value moves between registers.  Wider bands keep queues of 5 rows and
    // shift their rings (register budget).
    static constexpr bool FIX = M <= 4;
    // queue slots: with FIX a divisor of the unroll W.  Three slots (rows
    // r-1, r, r+1) prefetch one row ahead, which covers the HBM latency at
    // a row's duration (thousands of cycles) and keeps the registers for
    // occupancy.
    static constexpr int QN = FIX ? (W % 3 == 0 ? 3 : W) : 5;
    static constexpr int PF = QN - 2;          // field rows prefetched ahead
    static constexpr int UN = FIX ? W : QN;    // row-loop unroll (a multiple of QN)
    static_assert(UN % QN == 0, "queue slots must divide the unroll");
    static constexpr int RS = FIX ? W : W;     // F ring slots
};

int stream_strips(int ng, int m, bool ref) {
    const int so = 64 - 2 * (ref ? m + 1 : m);
    return (ng + so - 1) / so;
}

// a row of zeros the streaming kernel reads for source rows outside the box
__device__ float2 g_zero_row[4096];

// A u for u and the sum of its two neighbours, A = tridiag(-i a, 1 + 2 i a, -i a)
__device__ __forceinline__ double2 adi_apply(double2 u, double2 nb, double a) {
    return d2(u.x - 2.0 * a * u.y + a * nb.y, u.y + 2.0 * a * u.x - a * nb.x);
}
// u + D^-1 rho, D = 1 + 2 i a, with dn = 1 / (1 + 4 a^2)
__device__ __forceinline__ double2 jacobi_update(double2 u, double2 rho, double a, double dn) {
    return d2(u.x + (rho.x + 2.0 * a * rho.y) * dn, u.y + (rho.y - 2.0 * a * rho.x) * dn);
}

// strip `strip` of beam slice ii on one wavefront; lrow, lu1: [2][64] LDS rows of its own
template <int M, bool REF>
__device__ void adi_stream_body(const MixArrays &a, const int ii, const int strip, float2 (*lrow)[64],
                                float2 (*lu1)[64]) {
    using G = StreamGeom<M, REF>;
    constexpr int W = G::W, H = G::H, SO = G::SO, PF = G::PF, UN = G::UN;
    constexpr bool FIX = G::FIX;
    const int lane = threadIdx.x & 63;
    const MixStep &S = a.steps[a.st->istep];
    const MixBand bi = a.binfo[S.band];
    const float2 *band = a.bands + bi.off;
    const double ad = bi.a;
    const float af = (float)bi.a;
    const double dn = 1.0 / (1.0 + 4.0 * ad * ad);
    const int ng = a.ng, ns = a.ns;
    const int nstrip = (ng + SO - 1) / SO;
    const int x = strip * SO - H + lane;
    const bool xin = x >= 0 && x < ng;
    const bool bandl = lane >= M && lane < 64 - M;    // lanes with the band's whole support
    const bool outl = lane >= H && lane < 64 - H && xin;
    // every band lane of the strip on a column that shares the central row (uniform)
    const bool xsym = strip * SO - H + M >= bi.klo && strip * SO - H + 63 - M <= bi.khi;
    const long long fs = (long long)((ii + a.st->first) % ns);
    const int cur = a.st->cur;
    const float2 *u = a.field[cur] + fs * ng * ng;
    float2 *uo = a.field[cur ^ 1] + fs * ng * ng;
    const float2 *Sp = a.src + (long long)ii * ng * ng;
    const int4 bx = a.box[ii];
    const bool sxin = x >= bx.x && x <= bx.z;
    // this column's band row (edge strips only)
    float2 bxr[W];
#pragma unroll
    for (int m = 0; m < W; m++) bxr[m] = (xin && !xsym) ? band[(long long)x * W + m] : f2(0, 0);
    const float2 *bc = band + (long long)((bi.klo + bi.khi) / 2) * W;
    float2 bcr[M + 1];  // the central row's right half b_M .. b_2M (symmetric)
#pragma unroll
    for (int m = 0; m <= M; m++) bcr[m] = bc[M + m];
    // field rows q[k] = u(r - 1 + k) and source rows sq[k] = S(r + k - 1),
    // k = 0..PF+1, prefetched in the order they are consumed (the memory
    // counter retires loads in issue order)
    // Loads are unconditional from clamped (valid) addresses so they stay in
    // flight; the grid and source-box masks are applied when a row is used.
    const int xcl = min(max(x, 0), ng - 1);
    const float2 *ucol = u + xcl, *scol = Sp + xcl;
    auto uload = [&](int rr) -> float2 { return ucol[(long long)min(rr, ng - 1) * ng]; };
    // source rows outside the box read the zero row instead (no HBM traffic)
    auto sload = [&](int rr) -> float2 {
        const bool in = sxin && rr >= bx.y && rr <= bx.w;
        return *(in ? scol + (long long)rr * ng : g_zero_row + xcl);
    };
    // queues of QN = PF+2 rows, indexed modulo QN: the row loop is unrolled
    // by UN, a multiple of QN, so every slot is a fixed register (no register
    // moves of loads in flight).  Slot (r + 1 + k) % QN holds row r + k,
    // k = -1..PF.
    constexpr int QN = G::QN;
    float2 q[QN], sq[QN];
    q[0] = f2(0, 0);
    sq[0] = f2(0, 0);
#pragma unroll
    for (int k = 1; k < QN; k++) {
        q[k] = uload(k - 1);
        sq[k] = sload(k - 1);
    }
    float2 ring[W];
#pragma unroll
    for (int m = 0; m < W; m++) ring[m] = f2(0, 0);
    // REF: the FP32 remainders of F for rows r-M-1 .. r (FIX: W slots, row
    // r in slot r mod W), and u'' of the two rows before the band output
    constexpr int NL = REF ? (FIX ? W : M + 2) : 1;
    float2 lo[NL];
#pragma unroll
    for (int m = 0; m < NL; m++) lo[m] = f2(0, 0);
    float2 p1 = f2(0, 0), p2 = f2(0, 0);
    const double shift = -0.5 * (double)(ng - 1);
    const int icen = (ng / 2) * ng + (ng / 2);
    double dg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    auto emit = [&](int y, float2 o) {
        if (outl) {
            uo[(long long)y * ng + x] = o;
            const double ox = o.x, oy = o.y;
            const double wei = ox * ox + oy * oy;
            const double dx = (double)x + shift, dy = (double)y + shift;
            dg[0] += wei;
            dg[1] += dx * wei;
            dg[2] += dx * dx * wei;
            dg[3] += dy * wei;
            dg[4] += dy * dy * wei;
            dg[5] += ox;
            dg[6] += oy;
            if (y * ng + x == icen) dg[7] = wei;
        }
    };
    const int nrow = ng + M + (REF ? 1 : 0);
    for (int r0 = 0; r0 < nrow; r0 += UN) {
        // j = jj QN + jq: the queue slot of row r0 + j is a function of jq
        // alone (a remainder of the flat index here keeps the loop from
        // unrolling)
#pragma unroll
        for (int jj = 0; jj < UN / QN; jj++)
#pragma unroll
        for (int jq = 0; jq < QN; jq++) {
            const int j = jj * QN + jq;
            const int r = r0 + j;
            if (r >= nrow) break;
            // slots: row r-1 -> j, row r -> (j+1)%QN, row r+1 -> (j+2)%QN
            float2 Fr = f2(0, 0), Flo = f2(0, 0);
            if (r < ng) {
                const float2 sv = sq[(jq + 1) % QN];
                const float2 uc = q[(jq + 1) % QN];
                const float2 un = (r + 1 < ng) ? q[(jq + 2) % QN] : f2(0, 0);
                const float2 ul = (r > 0) ? q[jq] : f2(0, 0);
                double2 E64 = d2(0, 0);
                float2 E = f2(0, 0);
                if constexpr (REF) {
                    if (xin) {
                        const double sx = (double)un.x - 2.0 * (double)uc.x + (double)ul.x;
                        const double sy = (double)un.y - 2.0 * (double)uc.y + (double)ul.y;
                        E64 = d2((double)sv.x + (double)uc.x - ad * sy, (double)sv.y + (double)uc.y + ad * sx);
                    }
                    E = tof(E64);
                } else {
                    const float sx = un.x - 2.f * uc.x + ul.x, sy = un.y - 2.f * uc.y + ul.y;
                    E = xin ? f2(sv.x + uc.x - af * sy, sv.y + uc.y + af * sx) : f2(0, 0);
                }
                float2 *lr = lrow[j & 1];
                lr[lane] = E;
                // one wavefront per workgroup: LDS operations of a wave complete in
                // issue order, so the row is visible to the reads below without a
                // hardware barrier; a workgroup barrier would also wait for the
                // prefetched field rows (its fence drains vmcnt).  Compiler ordering only.
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                float2 up = f2(0, 0);
                if (bandl && xin) {
                    if (xsym) {
                        up = fmul(bcr[0], lr[lane]);
#pragma unroll
                        for (int k = 1; k <= M; k++) up = ffma(bcr[k], fadd(lr[lane - k], lr[lane + k]), up);
                    } else {
#pragma unroll
                        for (int m = 0; m < W; m++) up = ffma(bxr[m], lr[lane + m - M], up);
                    }
                }
                if constexpr (REF) {
                    float2 *lu = lu1[j & 1];
                    lu[lane] = up;  // zero off the grid and on the halo lanes' far side
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const float2 nl = lu[lane > 0 ? lane - 1 : 0], nr = lu[lane < 63 ? lane + 1 : 63];
                    const double2 u1 = tod(up);
                    const double2 nb = d2((double)nl.x + (double)nr.x, (double)nl.y + (double)nr.y);
                    const double2 au = adi_apply(u1, nb, ad);
                    const double2 upd = jacobi_update(u1, d2(E64.x - au.x, E64.y - au.y), ad, dn);
                    if (xin) {
                        const double2 F64 = d2((double)sv.x + 2.0 * upd.x - E64.x, (double)sv.y + 2.0 * upd.y - E64.y);
                        Fr = tof(F64);
                        Flo = f2((float)(F64.x - (double)Fr.x), (float)(F64.y - (double)Fr.y));
                    }
                } else {
                    Fr = xin ? f2(sv.x + 2.f * up.x - E.x, sv.y + 2.f * up.y - E.y) : f2(0, 0);
                }
                // slot j (row r-1) is free: it takes row r + PF + 1, first used
                // PF rows later
                q[jq] = uload(r + PF + 1);
                sq[jq] = sload(r + PF + 1);
            }
            // F of rows r - W + 1 .. r: window index k = 0..W-1 (k = W-1 the
            // newest row).  FIX: row r sits in slot r mod W = j, index k in
            // slot (j + 1 + k) mod W; otherwise the ring shifts down by one.
            if constexpr (FIX) {
                ring[j % W] = Fr;
                if constexpr (REF) lo[j % W] = Flo;
            } else {
#pragma unroll
                for (int m = 0; m < W - 1; m++) ring[m] = ring[m + 1];
                ring[W - 1] = Fr;
                if constexpr (REF) {
#pragma unroll
                    for (int m = 0; m < NL - 1; m++) lo[m] = lo[m + 1];
                    lo[NL - 1] = Flo;
                }
            }
            float2 rw[W];  // the window in index order (register renaming once unrolled)
#pragma unroll
            for (int k = 0; k < W; k++) {
                if constexpr (FIX) rw[k] = ring[(j + 1 + k) % W];
                else rw[k] = ring[k];
            }
            const int y = r - M;
            float2 o = f2(0, 0);
            if (y >= 0 && y < ng) {
                if (y >= bi.klo && y <= bi.khi) {
                    o = fmul(bcr[0], rw[M]);
#pragma unroll
                    for (int k = 1; k <= M; k++) o = ffma(bcr[k], fadd(rw[M - k], rw[M + k]), o);
                } else {
                    const float2 *by = band + (long long)y * W;
#pragma unroll
                    for (int m = 0; m < W; m++) o = ffma(by[m], rw[m], o);
                }
            }
            if constexpr (REF) {
                // row yf = y - 1 with u''(yf - 1) = p2, u''(yf) = p1, u''(yf + 1) = o;
                // F(yf) = window index M - 1 plus its remainder
                const int yf = y - 1;
                if (yf >= 0 && yf < ng) {
                    const float2 fh = rw[M - 1];
                    float2 fl;
                    if constexpr (FIX) fl = lo[(j + W - M - 1) % W];
                    else fl = lo[0];
                    const double2 F64 = d2((double)fh.x + (double)fl.x, (double)fh.y + (double)fl.y);
                    const double2 u2 = tod(p1);
                    const double2 nb = d2((double)p2.x + (double)o.x, (double)p2.y + (double)o.y);
                    const double2 au = adi_apply(u2, nb, ad);
                    emit(yf, tof(jacobi_update(u2, d2(F64.x - au.x, F64.y - au.y), ad, dn)));
                }
                p2 = p1;
                p1 = o;
            } else {
                if (y >= 0) emit(y, o);
            }
        }
    }
    // strip sums (wave butterfly, fixed order)
#pragma unroll
    for (int k = 0; k < 8; k++)
        for (int o = 32; o > 0; o >>= 1) dg[k] += __shfl_xor(dg[k], o);
    if (lane == 0) {
        double *dst = a.tilediag + (fs * nstrip + strip) * 8;
        for (int k = 0; k < 8; k++) dst[k] = dg[k];
    }
}

template <int M, bool REF>
__global__ void __launch_bounds__(64) k_mix_adi_stream(MixArrays a) {
    __shared__ float2 lrow[2][64];
    __shared__ float2 lu1[2][64];
    constexpr int SO = StreamGeom<M, REF>::SO;
    const int nstrip = (a.ng + SO - 1) / SO;
    const int b = blockIdx.x;
    const int ii = (b / (8 * nstrip)) * 8 + (b & 7);   // XCD-aware: a slice's strips on one XCD
    const int strip = (b % (8 * nstrip)) >> 3;
    if (ii >= a.ns) return;
    adi_stream_body<M, REF>(a, ii, strip, lrow, lu1);
}

// ---------------------------------------------------------------------------
// Whole-width streaming ADI: the row recurrence of adi_stream_body with one
// workgroup of NW = blockDim / 64 wavefronts covering every column of a
// slice (one lane per column), so no column is computed twice; the strips
// above recompute a 2H-column halo (320 lanes for 255 columns at M = 4 with
// the correction).  The E and u1 rows pass between the wavefronts through
// LDS with one workgroup barrier per exchange whose fences cover LDS only,
// so the field rows prefetched from HBM stay in flight across it.  Cells off
// the grid hold zeros in the LDS rows: the Dirichlet boundary of A.
//
// Column map: wavefront 0 takes the first and the last 32 columns, whose
// band rows differ from the central one; wavefronts 1.. take the columns in
// between and use the symmetric central row when all of theirs lie in
// [klo, khi].  Lanes left over sit on column ng (off the grid, zero).
//
// Buffers: the E row is double-buffered by row parity.  Without the
// correction a wavefront may write row r+1 while another still reads row r
// (one barrier per row).  With it, a second barrier between the E reads and
// the u1 reads separates every write of a row from the reads of the row
// before, so u1 needs one buffer.
//
// Dynamic LDS: E rows [2][NC + 2M] float2, u1 row [NC + 2] float2 (with the
// correction), per-column band rows [W][NC] float2, wavefront sums [NW][8].
// ---------------------------------------------------------------------------
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int M, bool REF>
size_t wide_lds_bytes(int nw) {
    const int nc = 64 * nw;
    return (size_t)2 * (nc + 2 * M) * sizeof(float2) + (REF ? (size_t)(nc + 2) * sizeof(double2) : 0) +
           (size_t)(2 * M + 1) * nc * sizeof(float2) + (size_t)nw * 8 * sizeof(double);
}

// three wavefronts per SIMD where the registers allow it without spilling
template <int M, bool REF>
__global__ void __launch_bounds__(512, (REF && M == 9) ? 2 : 3) k_mix_adi_wide(MixArrays a) {
    using G = StreamGeom<M, REF>;
    constexpr int W = G::W, QN = G::QN, PF = G::PF, UN = G::UN;
    constexpr bool FIX = G::FIX;
    extern __shared__ double2 wide_lds[];
    const int nt = blockDim.x, nw = nt >> 6, nc = nt;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    float2 *lrow0 = (float2 *)wide_lds;                    // [2][nc + 2M]
    const int lp = nc + 2 * M;
    double2 *lu = (double2 *)(lrow0 + 2 * lp);              // [nc + 2] u1 row (FP32 values, widened)
    float2 *lband = (float2 *)(lu + (REF ? nc + 2 : 0));    // [W][nc]
    double *wsum = (double *)(lband + W * nc);              // [nw][8]
    const int ii = blockIdx.x;
    if (ii >= a.ns) return;
    const MixStep &S = a.steps[a.st->istep];
    const MixBand bi = a.binfo[S.band];
    const float2 *band = a.bands + bi.off;
    const double ad = bi.a;
    const float af = (float)bi.a;
    const double dn = 1.0 / (1.0 + 4.0 * ad * ad);
    const int ng = a.ng, ns = a.ns;
    // this lane's column
    int x;
    if (ng <= 64) x = lane < ng ? lane : ng;
    else if (w == 0) x = lane < 32 ? lane : ng - 64 + lane;
    else {
        x = 32 + 64 * (w - 1) + lane;
        if (x >= ng - 32) x = ng;
    }
    const bool xin = x < ng;
    // every grid column of this wavefront shares the central band row (uniform)
    bool wsym;
    if (ng <= 64 || w == 0) wsym = bi.klo <= 0 && bi.khi >= ng - 1;
    else {
        const int xlo = 32 + 64 * (w - 1), xhi = min(xlo + 63, ng - 33);
        wsym = xlo > xhi || (xlo >= bi.klo && xhi <= bi.khi);
    }
    // zero rows (the pads stay zero: only lanes on the grid write non-zero values)
    for (int k = t; k < 2 * lp; k += nt) lrow0[k] = f2(0, 0);
    if constexpr (REF)
        for (int k = t; k < nc + 2; k += nt) lu[k] = d2(0, 0);
    if (!wsym)
        for (int m = 0; m < W; m++) lband[m * nc + t] = xin ? band[(long long)x * W + m] : f2(0, 0);
    const float2 *bc = band + (long long)((bi.klo + bi.khi) / 2) * W;
    float2 bcr[M + 1];  // the central row's right half b_M .. b_2M (symmetric)
#pragma unroll
    for (int m = 0; m <= M; m++) bcr[m] = bc[M + m];
    const long long fs = (long long)((ii + a.st->first) % ns);
    const int cur = a.st->cur;
    const float2 *u = a.field[cur] + fs * ng * ng;
    float2 *uo = a.field[cur ^ 1] + fs * ng * ng;
    const float2 *Sp = a.src + (long long)ii * ng * ng;
    const int4 bx = a.box[ii];
    const bool sxin = xin && x >= bx.x && x <= bx.z;
    const int xcl = min(x, ng - 1);
    const float2 *ucol = u + xcl, *scol = Sp + xcl;
    auto uload = [&](int rr) -> float2 { return ucol[(long long)min(rr, ng - 1) * ng]; };
    auto sload = [&](int rr) -> float2 {
        const bool in = sxin && rr >= bx.y && rr <= bx.w;
        return *(in ? scol + (long long)rr * ng : g_zero_row + xcl);
    };
    float2 q[QN], sq[QN];
    q[0] = f2(0, 0);
    sq[0] = f2(0, 0);
#pragma unroll
    for (int k = 1; k < QN; k++) {
        q[k] = uload(k - 1);
        sq[k] = sload(k - 1);
    }
    float2 ring[W];
#pragma unroll
    for (int m = 0; m < W; m++) ring[m] = f2(0, 0);
    constexpr int NL = REF ? (FIX ? W : M + 2) : 1;
    float2 lo[NL];
#pragma unroll
    for (int m = 0; m < NL; m++) lo[m] = f2(0, 0);
    float2 p1 = f2(0, 0), p2 = f2(0, 0);
    // REF: the field rows r-1, r (and r+1 while row r is formed) widened to
    // FP64 once each, as they enter the three-row window
    double2 d_l = d2(0, 0), d_c = tod(q[1]);
    // field sums of this column: x is fixed, so the x moments are dx, dx^2
    // times the column's power, applied at the end
    double s_w = 0, s_yw = 0, s_yyw = 0, s_re = 0, s_im = 0, s_cen = 0;
    const bool xcen = x == ng / 2;
    const double shift = -0.5 * (double)(ng - 1);
    // output row y: the stored FP32 value o, the sums from o64 (the value
    // before rounding with the correction)
    auto emit = [&](int y, float2 o, double2 o64) {
        if (xin) {
            uo[(long long)y * ng + x] = o;
            const double ox = o64.x, oy = o64.y;
            const double wei = ox * ox + oy * oy;
            const double dy = (double)y + shift;
            s_w += wei;
            s_yw += dy * wei;
            s_yyw += dy * dy * wei;
            s_re += ox;
            s_im += oy;
            if (xcen && y == ng / 2) s_cen = wei;
        }
    };
    lds_barrier();
    float2 *lx = lrow0 + x + M;  // this column in E row 0; row 1 is lp further
    const int nrow = ng + M + (REF ? 1 : 0);
    for (int r0 = 0; r0 < nrow; r0 += UN) {
        // j = jj QN + jq: the queue slot of row r0 + j is a function of jq
        // alone (a remainder of the flat index here keeps the loop from
        // unrolling)
#pragma unroll
        for (int jj = 0; jj < UN / QN; jj++)
#pragma unroll
        for (int jq = 0; jq < QN; jq++) {
            const int j = jj * QN + jq;
            const int r = r0 + j;
            if (r >= nrow) break;
            float2 Fr = f2(0, 0), Flo = f2(0, 0);
            if (r < ng) {
                const float2 sv = sq[(jq + 1) % QN];
                const float2 uc = q[(jq + 1) % QN];
                const float2 un = (r + 1 < ng) ? q[(jq + 2) % QN] : f2(0, 0);
                const float2 ul = (r > 0) ? q[jq] : f2(0, 0);
                double2 E64 = d2(0, 0);
                float2 E = f2(0, 0);
                double2 sv64 = d2(0, 0);
                if constexpr (REF) {
                    const double2 d_n = (r + 1 < ng) ? tod(un) : d2(0, 0);
                    if (xin) {
                        sv64 = tod(sv);
                        const double sx = fma(-2.0, d_c.x, d_n.x) + d_l.x;
                        const double sy = fma(-2.0, d_c.y, d_n.y) + d_l.y;
                        E64 = d2(fma(-ad, sy, sv64.x + d_c.x), fma(ad, sx, sv64.y + d_c.y));
                    }
                    d_l = d_c;
                    d_c = d_n;
                    E = tof(E64);
                } else {
                    const float sx = un.x - 2.f * uc.x + ul.x, sy = un.y - 2.f * uc.y + ul.y;
                    E = xin ? f2(sv.x + uc.x - af * sy, sv.y + uc.y + af * sx) : f2(0, 0);
                }
                float2 *lr = lx + (r & 1) * lp;
                lr[0] = E;
                lds_barrier();
                float2 up = f2(0, 0), up1 = f2(0, 0);  // two partial sums (even, odd terms)
                if (xin) {
                    if (wsym) {
                        up = fmul(bcr[0], lr[0]);
#pragma unroll
                        for (int k = 1; k <= M; k++) {
                            if (k & 1) up1 = ffma(bcr[k], fadd(lr[-k], lr[k]), up1);
                            else up = ffma(bcr[k], fadd(lr[-k], lr[k]), up);
                        }
                    } else {
#pragma unroll
                        for (int m = 0; m < W; m++) {
                            if (m & 1) up1 = ffma(lband[m * nc + t], lr[m - M], up1);
                            else up = ffma(lband[m * nc + t], lr[m - M], up);
                        }
                    }
                    up = fadd(up, up1);
                }
                if constexpr (REF) {
                    // u1 goes to the row widened once, not by each neighbour
                    const double2 u1 = tod(up);
                    lu[x + 1] = u1;  // zero off the grid
                    lds_barrier();
                    const double2 nl = lu[x], nr = lu[x + 2];
                    const double2 au = adi_apply(u1, d2(nl.x + nr.x, nl.y + nr.y), ad);
                    const double2 upd = jacobi_update(u1, d2(E64.x - au.x, E64.y - au.y), ad, dn);
                    if (xin) {
                        const double2 F64 = d2(fma(2.0, upd.x, sv64.x) - E64.x, fma(2.0, upd.y, sv64.y) - E64.y);
                        Fr = tof(F64);
                        Flo = f2((float)(F64.x - (double)Fr.x), (float)(F64.y - (double)Fr.y));
                    }
                } else {
                    Fr = xin ? f2(sv.x + 2.f * up.x - E.x, sv.y + 2.f * up.y - E.y) : f2(0, 0);
                }
                q[jq] = uload(r + PF + 1);
                sq[jq] = sload(r + PF + 1);
            }
            if constexpr (FIX) {
                ring[j % W] = Fr;
                if constexpr (REF) lo[j % W] = Flo;
            } else {
#pragma unroll
                for (int m = 0; m < W - 1; m++) ring[m] = ring[m + 1];
                ring[W - 1] = Fr;
                if constexpr (REF) {
#pragma unroll
                    for (int m = 0; m < NL - 1; m++) lo[m] = lo[m + 1];
                    lo[NL - 1] = Flo;
                }
            }
            float2 rw[W];
#pragma unroll
            for (int k = 0; k < W; k++) {
                if constexpr (FIX) rw[k] = ring[(j + 1 + k) % W];
                else rw[k] = ring[k];
            }
            const int y = r - M;
            float2 o = f2(0, 0), o1 = f2(0, 0);
            if (y >= 0 && y < ng) {
                if (y >= bi.klo && y <= bi.khi) {
                    o = fmul(bcr[0], rw[M]);
#pragma unroll
                    for (int k = 1; k <= M; k++) {
                        if (k & 1) o1 = ffma(bcr[k], fadd(rw[M - k], rw[M + k]), o1);
                        else o = ffma(bcr[k], fadd(rw[M - k], rw[M + k]), o);
                    }
                } else {
                    const float2 *by = band + (long long)y * W;
#pragma unroll
                    for (int m = 0; m < W; m++) {
                        if (m & 1) o1 = ffma(by[m], rw[m], o1);
                        else o = ffma(by[m], rw[m], o);
                    }
                }
                o = fadd(o, o1);
            }
            if constexpr (REF) {
                const int yf = y - 1;
                if (yf >= 0 && yf < ng) {
                    const float2 fh = rw[M - 1];
                    float2 fl;
                    if constexpr (FIX) fl = lo[(j + W - M - 1) % W];
                    else fl = lo[0];
                    const double2 F64 = d2((double)fh.x + (double)fl.x, (double)fh.y + (double)fl.y);
                    const double2 u2 = tod(p1);
                    const double2 nb = d2((double)p2.x + (double)o.x, (double)p2.y + (double)o.y);
                    const double2 au = adi_apply(u2, nb, ad);
                    const double2 res = jacobi_update(u2, d2(F64.x - au.x, F64.y - au.y), ad, dn);
                    emit(yf, tof(res), res);
                }
                p2 = p1;
                p1 = o;
            } else {
                if (y >= 0) emit(y, o, tod(o));
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
        for (int q2 = 1; q2 < nw; q2++) v += wsum[q2 * 8 + t];
        a.tilediag[fs * 8 + t] = v;
    }
}

// ---------------------------------------------------------------------------
// slippage on the device (Control::applySlippage, Control.cpp:158-276): the
// ring offset, the slices to zero, the live field buffer and the step index.
// ---------------------------------------------------------------------------
__device__ void slip_body(const MixArrays &a) {
    MixState &st = *a.st;
    const MixStep &S = a.steps[st.istep];
    const int ns = a.ns;
    st.nzero = 0;
    st.cur ^= 1;  // the ADI of this step wrote the other buffer
    if (a.timerun) {
        st.accuslip += S.slip;
        int direction = 1;
        while (fabs(st.accuslip) > (a.sample * 0.8)) {
            if (st.accuslip < 0) direction = -1;
            st.accuslip -= a.sample * direction;
            int last = (st.first + ns - 1) % ns;
            if (direction < 0) last = (last + 1) % ns;
            if (!a.periodic && st.nzero < kMixMaxZero) st.zero[st.nzero++] = last;
            st.first = last;
            if (direction < 0) st.first = (last + 1) % ns;
        }
    }
    st.rec_step = st.istep;
    st.istep = st.istep + 1;
    st.seq = st.seq + 1;
}

__global__ void k_mix_slip(MixArrays a) {
    if (threadIdx.x == 0) slip_body(a);
}

__global__ void __launch_bounds__(256) k_mix_zero(MixArrays a) {
    const int z = blockIdx.y;
    if (z >= a.st->nzero) return;
    const long long cells = (long long)a.ng * a.ng;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < cells) a.field[a.st->cur][(long long)a.st->zero[z] * cells + i] = f2(0, 0);
}

// Record of the completed step per beam slice: beam sums and the field sums
// of its storage slice (zero for slices slippage has just cleared).
__device__ void record_body(const MixArrays &a, const int is, int ntiles) {
    const MixState &st = *a.st;
    const int fs = (is + st.first) % a.ns;
    bool zeroed = false;
    for (int k = 0; k < st.nzero; k++) zeroed |= (st.zero[k] == fs);
    double f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!zeroed)
        for (int k = 0; k < ntiles; k++)
            for (int q = 0; q < 8; q++) f[q] += a.tilediag[((long long)fs * ntiles + k) * 8 + q];
    double *o = a.rec + ((long long)(st.rec_step % a.nrec) * a.ns + is) * kMixRecord;
    for (int q = 0; q < 8; q++) o[q] = a.beamsum[8 * is + q];
    for (int q = 0; q < 8; q++) o[8 + q] = f[q];
}

__global__ void __launch_bounds__(64) k_mix_record(MixArrays a, int ntiles) {
    const int is = blockIdx.x * 64 + threadIdx.x;
    if (is < a.ns) record_body(a, is, ntiles);
}

// Standalone sums for a diagnostics call outside a run.
__global__ void __launch_bounds__(256) k_mix_diag_beam(MixArrays a, double *out) {
    __shared__ double lds[8 * 256];
    const int is = blockIdx.x;
    const long long b0 = a.off[is], b1 = a.off[is + 1];
    const float ph = wrap_phase(a.base[is]);
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long long i = b0 + threadIdx.x; i < b1; i += 256) {
        const double gd = a.gref + (double)a.dg[i];
        const double x = a.x[i], y = a.y[i];
        v[0] += gd;
        v[1] += gd * gd;
        v[2] += x;
        v[3] += x * x;
        v[4] += y;
        v[5] += y * y;
        float s, c;
        __sincosf(ph + a.dth[i], &s, &c);
        v[6] += c;
        v[7] += s;
    }
    bsum<256, 8>(v, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; k++) out[8 * is + k] = v[k];
}

__global__ void __launch_bounds__(256) k_mix_diag_field(MixArrays a, double *out, int next) {
    __shared__ double lds[8 * 256];
    const int is0 = blockIdx.x, ng = a.ng;
    const float2 *f = a.field[a.st->cur ^ (next ? 1 : 0)] + (long long)is0 * ng * ng;
    const double shift = -0.5 * (double)(ng - 1);
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < ng * ng; i += 256) {
        const int iy = i / ng, ix = i % ng;
        const double dy = (double)iy + shift, dx = (double)ix + shift;
        const double lx = f[i].x, ly = f[i].y;
        const double wei = lx * lx + ly * ly;
        v[0] += wei;
        v[1] += dx * wei;
        v[2] += dx * dx * wei;
        v[3] += dy * wei;
        v[4] += dy * dy * wei;
        v[5] += lx;
        v[6] += ly;
    }
    bsum<256, 8>(v, lds);
    if (threadIdx.x == 0) {
        const int ic = (ng / 2) * ng + (ng / 2);
        const double cx = f[ic].x, cy = f[ic].y;
        v[7] = cx * cx + cy * cy;
        for (int k = 0; k < 8; k++) out[8 * is0 + k] = v[k];
    }
}

// Sum of |S|^2 over the source box of each beam slice (the self-energy term
// of a field step's energy balance).
__global__ void __launch_bounds__(256) k_mix_src_sums(MixArrays a, double *out) {
    __shared__ double red[256];
    const int is = blockIdx.x, ng = a.ng;
    const int4 bx = a.box[is];
    const float2 *S = a.src + (long long)is * ng * ng;
    const int w = bx.z - bx.x + 1, h = bx.w - bx.y + 1;
    double v = 0;
    if (w > 0 && h > 0)
        for (int k = threadIdx.x; k < w * h; k += 256) {
            const float2 q = S[(bx.y + k / w) * ng + bx.x + k % w];
            v += (double)q.x * q.x + (double)q.y * q.y;
        }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[is] = red[0];
}

// Sorting::localSort target slice (src/Util/Sorting.cpp:96) on the absolute
// phase base + dth.
__global__ void __launch_bounds__(256) k_mix_sort_index(MixArrays a, double invslen, int *atar,
                                                        unsigned long long *moved) {
    const int is = blockIdx.x;
    const long long b0 = a.off[is], b1 = a.off[is + 1];
    const double base = a.base[is];
    for (long long i0 = b0; i0 < b1; i0 += 256) {
        const long long i = i0 + threadIdx.x;
        int m = 0;
        if (i < b1) {
            const int at = (int)floor((base + (double)a.dth[i]) * invslen);
            atar[i] = at;
            m = (at != 0);
        }
        const unsigned long long bal = __ballot(m);
        if ((threadIdx.x & 63) == 0 && bal) atomicAdd(moved, (unsigned long long)__popcll(bal));
    }
}

// ---------------------------------------------------------------------------
// The whole step in one launch: every phase of launch_mix_step on a grid of
// co-resident workgroups (a cooperative launch, so every workgroup is
// resident and reaches every barrier), phases separated by grid barriers:
//   front (all slices) | radial scan | main (all slices) | ADI (all strips)
//   | slippage | zeroing + records.
// Workgroups of kPerNT threads walk the slices; in the ADI phase every
// wavefront walks (slice, strip) items, the strips of a slice on the
// workgroups of one XCD (workgroups are dispatched to the XCDs round-robin,
// so block b runs on XCD b % 8: locality only, not correctness).  The
// dynamic LDS region is the main kernel's; front and ADI carve their
// scratch from it.  Every loop is uniform per workgroup and nothing returns
// early, so each workgroup meets each grid barrier once.
// ---------------------------------------------------------------------------
constexpr int kPerNT = 384;

template <int M, bool REF>
__global__ void __launch_bounds__(kPerNT) k_mix_step_persistent(MixArrays a, int sc, int moments, int maxzero) {
    extern __shared__ float2 dyn[];
    cg::grid_group grid = cg::this_grid();
    const int G = (int)gridDim.x, b = (int)blockIdx.x, t = (int)threadIdx.x, w = t >> 6;
    const int ns = a.ns;
    for (int is = b; is < ns; is += G) front_step<kPerNT>(a, is, moments, sc, (double *)dyn);
    grid.sync();
    if (sc && b == 0) scan_body(a, (double *)dyn);  // uniform per workgroup
    grid.sync();
    for (int is = b; is < ns; is += G) main_body<false, kPerNT, kMixDyn>(a, is, sc, dyn);
    grid.sync();
    constexpr int SO = StreamGeom<M, REF>::SO;
    const int nstrip = (a.ng + SO - 1) / SO;
    {
        float2(*lrow)[64] = (float2(*)[64])dyn + 4 * w;  // 4 rows of 64 per wavefront
        float2(*lu1)[64] = lrow + 2;
        const int x = b & 7, nbx = (G - x + 7) / 8, j = b >> 3;
        const int nsx = (ns - x + 7) / 8;                // slices x, x + 8, ... of this XCD
        for (int L = j * (kPerNT / 64) + w; L < nsx * nstrip; L += nbx * (kPerNT / 64))
            adi_stream_body<M, REF>(a, x + 8 * (L / nstrip), L % nstrip, lrow, lu1);
    }
    grid.sync();
    if (b == 0 && t == 0) slip_body(a);
    grid.sync();
    const long long cells = (long long)a.ng * a.ng;
    const long long nzc = (long long)min(a.st->nzero, maxzero) * cells;
    for (long long e = (long long)b * kPerNT + t; e < nzc; e += (long long)G * kPerNT)
        a.field[a.st->cur][(long long)a.st->zero[e / cells] * cells + e % cells] = f2(0, 0);
    for (int is = b * kPerNT + t; is < ns; is += G * kPerNT) record_body(a, is, nstrip);
}

template <int M, bool REF>
int set_adi_attr() {
    using Gm = MixGeom<M, REF>;
    if (Gm::lds_bytes > 160 * 1024) return 0;  // not launchable; never selected
    return hipFuncSetAttribute((const void *)k_mix_adi<M, REF>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)Gm::lds_bytes) ==
           hipSuccess;
}

template <int M, bool REF>
void launch_adi_t(hipStream_t s, const MixArrays &a) {
    using Gm = MixGeom<M, REF>;
    const int ntx = (a.ng + Gm::TS - 1) / Gm::TS;
    const unsigned nblocks = (unsigned)(((a.ns + 7) / 8) * 8) * (unsigned)(ntx * ntx);
    hipLaunchKernelGGL((k_mix_adi<M, REF>), dim3(nblocks), dim3(256), Gm::lds_bytes, s, a);
}

constexpr int kTmplM[] = {3, 4, 6, 9, 12};

}  // namespace

bool mix_band_supported(int m, int *tmpl_m) {
    for (int t : kTmplM)
        if (m <= t) {
            *tmpl_m = t;
            return true;
        }
    return false;
}

int mix_band_tiles(int ng) {
    const int ntx = (ng + 31) / 32;
    // the tile-diagnostic scratch also holds the strip sums of the streaming kernel
    return std::max(ntx * ntx, stream_strips(ng, 12, true));
}

size_t mix_band_lds(int tmpl_m, bool refine) {
#define G4_L(MM) \
    if (tmpl_m == MM) return refine ? MixGeom<MM, true>::lds_bytes : MixGeom<MM, false>::lds_bytes;
    G4_L(3) G4_L(4) G4_L(6) G4_L(9) G4_L(12)
#undef G4_L
    return 0;
}

template <bool EG, int NT, int DYN>
int set_main_attr() {
    return hipFuncSetAttribute((const void *)k_mix_main<EG, NT, DYN>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               DYN * (int)sizeof(float2)) == hipSuccess;
}

template <bool EG, int NT, int DYN>
void launch_main_t(hipStream_t s, const MixArrays &a, int sc) {
    hipLaunchKernelGGL((k_mix_main<EG, NT, DYN>), dim3(a.ns), dim3(NT), DYN * sizeof(float2), s, a, sc);
}

template <bool EG>
void launch_main(hipStream_t s, const MixArrays &a, int sc, int nt, int lds) {
    if (lds == 2 && nt != 512) {  // eight wavefronts' histograms need more than 32 KiB
        if (nt == 256) launch_main_t<EG, 256, kMixDynT>(s, a, sc);
        else launch_main_t<EG, 384, kMixDynT>(s, a, sc);
    } else if (lds >= 1) {
        if (nt == 256) launch_main_t<EG, 256, kMixDynS>(s, a, sc);
        else if (nt == 512) launch_main_t<EG, 512, kMixDynS>(s, a, sc);
        else launch_main_t<EG, 384, kMixDynS>(s, a, sc);
    } else {
        if (nt == 256) launch_main_t<EG, 256, kMixDyn>(s, a, sc);
        else if (nt == 512) launch_main_t<EG, 512, kMixDyn>(s, a, sc);
        else launch_main_t<EG, 384, kMixDyn>(s, a, sc);
    }
}

constexpr size_t kLdsMax = 160 * 1024;

// the whole-width kernel's LDS limit: what eight wavefronts need, at most
// the CU's 160 KiB (wider grids with wide bands then run the strips)
template <int M, bool REF>
int set_wide_attr() {
    return hipFuncSetAttribute((const void *)k_mix_adi_wide<M, REF>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)std::min(wide_lds_bytes<M, REF>(8), kLdsMax)) == hipSuccess;
}

size_t wide_lds_need(int tmpl_m, bool refine, int nw) {
#define G4_L(MM) \
    if (tmpl_m == MM) return refine ? wide_lds_bytes<MM, true>(nw) : wide_lds_bytes<MM, false>(nw);
    G4_L(3) G4_L(4) G4_L(6) G4_L(9) G4_L(12)
#undef G4_L
    return ~(size_t)0;
}

template <int M, bool REF>
int set_persistent_attr() {
    return hipFuncSetAttribute((const void *)k_mix_step_persistent<M, REF>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, kMixDyn * (int)sizeof(float2)) ==
           hipSuccess;
}

template <int M, bool REF>
int launch_persistent_t(hipStream_t s, const MixArrays &a, bool sc, bool moments, int maxzero) {
    const void *f = (const void *)k_mix_step_persistent<M, REF>;
    const size_t dynb = kMixDyn * sizeof(float2);
    int nb = 0, dev = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kPerNT, dynb) != hipSuccess || nb < 1) return -1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        return -1;
    MixArrays aa = a;
    int sci = sc ? 1 : 0, momi = moments ? 1 : 0, mz = maxzero;
    void *args[] = {&aa, &sci, &momi, &mz};
    // co-resident workgroups only: the launch fails (no hang) if they do not fit
    return hipLaunchCooperativeKernel(f, dim3((unsigned)(nb * ncu)), dim3(kPerNT), args, (unsigned)dynb, s) ==
                   hipSuccess
               ? 0
               : -1;
}

int launch_mix_step_persistent(hipStream_t s, const MixArrays &a, int tmpl_m, bool refine, bool sc, bool moments,
                               int maxzero, hipEvent_t *ev) {
    if (ev) (void)hipEventRecord(ev[0], s);
    int rc = -1;
#define G4_Q(MM) \
    if (tmpl_m == MM) rc = refine ? launch_persistent_t<MM, true>(s, a, sc, moments, maxzero) \
                                  : launch_persistent_t<MM, false>(s, a, sc, moments, maxzero);
    G4_Q(3) G4_Q(4) G4_Q(6) G4_Q(9) G4_Q(12)
#undef G4_Q
    if (ev) (void)hipEventRecord(ev[8], s);
    return rc;
}

int mix_prepare() {
    int ok = 1;
#define G4_M(EG, NT) ok &= set_main_attr<EG, NT, kMixDyn>(); ok &= set_main_attr<EG, NT, kMixDynS>();
    G4_M(true, 256) G4_M(true, 384) G4_M(true, 512) G4_M(false, 256) G4_M(false, 384) G4_M(false, 512)
    ok &= set_main_attr<true, 256, kMixDynT>() && set_main_attr<true, 384, kMixDynT>() &&
          set_main_attr<false, 256, kMixDynT>() && set_main_attr<false, 384, kMixDynT>();
#undef G4_M
#define G4_P(MM) ok &= set_adi_attr<MM, false>(); if (MixGeom<MM, true>::lds_bytes <= 160 * 1024) ok &= set_adi_attr<MM, true>(); \
    ok &= set_persistent_attr<MM, false>(); ok &= set_persistent_attr<MM, true>(); \
    ok &= set_wide_attr<MM, false>(); ok &= set_wide_attr<MM, true>();
    G4_P(3) G4_P(4) G4_P(6) G4_P(9) G4_P(12)
#undef G4_P
    return ok ? 0 : -1;
}

int mix_sc_mode_cells() { return kCelmCells; }

void launch_mix_beam(hipStream_t s, const MixArrays &a, bool sc, bool moments, hipEvent_t *ev, bool egsum,
                     int main_nt, int main_lds, bool front) {
    auto mark = [&](int k) {
        if (ev) (void)hipEventRecord(ev[k], s);
    };
    mark(0);
    if (front) hipLaunchKernelGGL(k_mix_front, dim3(a.ns), dim3(kFrontNT), 0, s, a, moments ? 1 : 0, sc ? 1 : 0);
    if (sc) hipLaunchKernelGGL(k_mix_scan, dim3(1), dim3(kScanNT), 0, s, a);
    mark(1);
    mark(2);
    if (egsum) launch_main<true>(s, a, sc ? 1 : 0, main_nt, main_lds);
    else launch_main<false>(s, a, sc ? 1 : 0, main_nt, main_lds);
    mark(3);
    mark(4);
    mark(5);
    mark(6);
}

int launch_mix_adi(hipStream_t s, const MixArrays &a, int tmpl_m, bool refine, int mode, int *ntiles) {
    // eight wavefronts cover 512 columns; the LDS of one workgroup at most 160 KiB
    if (mode == kAdiWide &&
        (a.ng > 512 || wide_lds_need(tmpl_m, refine, std::max(1, (a.ng + 63) / 64)) > kLdsMax))
        mode = kAdiStrips;
    const bool stream = mode != kAdiTile;
    if (refine && !stream && mix_band_lds(tmpl_m, true) > 160 * 1024) return -1;
    if (mode == kAdiWide) {
        // whole-width kernel: one workgroup of ceil(ng / 64) wavefronts per slice
        *ntiles = 1;
        const int nw = std::max(1, (a.ng + 63) / 64);
#define G4_W(MM)                                                                                          \
    else if (tmpl_m == MM) {                                                                              \
        if (refine) {                                                                                     \
            const size_t lds = wide_lds_bytes<MM, true>(nw);                                              \
            hipLaunchKernelGGL((k_mix_adi_wide<MM, true>), dim3(a.ns), dim3(64 * nw), lds, s, a);         \
        } else {                                                                                          \
            const size_t lds = wide_lds_bytes<MM, false>(nw);                                             \
            hipLaunchKernelGGL((k_mix_adi_wide<MM, false>), dim3(a.ns), dim3(64 * nw), lds, s, a);        \
        }                                                                                                 \
    }
        if (false) {
        }
        G4_W(3) G4_W(4) G4_W(6) G4_W(9) G4_W(12)
        else return -1;
#undef G4_W
        return 0;
    }
#define G4_A(MM)                                         \
    else if (tmpl_m == MM) {                             \
        if (refine) launch_adi_t<MM, true>(s, a);        \
        else launch_adi_t<MM, false>(s, a);              \
    }
    *ntiles = (a.ng + 31) / 32 * ((a.ng + 31) / 32);
    if (stream) {
        // streaming kernel: one wavefront per strip
        *ntiles = stream_strips(a.ng, tmpl_m, refine);
        const unsigned nblocks = (unsigned)(((a.ns + 7) / 8) * 8) * (unsigned)*ntiles;
#define G4_S(MM)                                                                                       \
    else if (tmpl_m == MM) {                                                                           \
        if (refine) hipLaunchKernelGGL((k_mix_adi_stream<MM, true>), dim3(nblocks), dim3(64), 0, s, a); \
        else hipLaunchKernelGGL((k_mix_adi_stream<MM, false>), dim3(nblocks), dim3(64), 0, s, a);      \
    }
        if (false) {
        }
        G4_S(3) G4_S(4) G4_S(6) G4_S(9) G4_S(12)
        else return -1;
#undef G4_S
    } else if (false) {
    }
    G4_A(3) G4_A(4) G4_A(6) G4_A(9) G4_A(12)
    else return -1;
#undef G4_A
    return 0;
}

void launch_mix_tail(hipStream_t s, const MixArrays &a, int maxzero, int ntiles, hipEvent_t *ev) {
    if (ev) (void)hipEventRecord(ev[7], s);
    hipLaunchKernelGGL(k_mix_slip, dim3(1), dim3(64), 0, s, a);
    if (maxzero > 0) {
        const long long cells = (long long)a.ng * a.ng;
        hipLaunchKernelGGL(k_mix_zero, dim3((unsigned)((cells + 255) / 256), (unsigned)maxzero),
                           dim3(256), 0, s, a);
    }
    hipLaunchKernelGGL(k_mix_record, dim3((a.ns + 63) / 64), dim3(64), 0, s, a, ntiles);
    if (ev) (void)hipEventRecord(ev[8], s);
}

int launch_mix_step(hipStream_t s, const MixArrays &a, int tmpl_m, bool refine, bool sc,
                    bool moments, int maxzero, hipEvent_t *ev, int adi_mode, int main_nt, int main_lds,
                    bool front) {
    launch_mix_beam(s, a, sc, moments, ev, false, main_nt, main_lds, front);
    int ntiles = 0;
    if (launch_mix_adi(s, a, tmpl_m, refine, adi_mode, &ntiles) != 0) return -1;
    launch_mix_tail(s, a, maxzero, ntiles, ev);
    return 0;
}

void launch_mix_src_sums(hipStream_t s, const MixArrays &a, double *out) {
    hipLaunchKernelGGL(k_mix_src_sums, dim3(a.ns), dim3(256), 0, s, a, out);
}

void launch_mix_field_sums(hipStream_t s, const MixArrays &a, int next, double *out8) {
    hipLaunchKernelGGL(k_mix_diag_field, dim3(a.ns), dim3(256), 0, s, a, out8, next);
}

// The auxiliary beam diagnostics and the on-axis field value per beam slice
// (the FP64 form's k_diag_aux on the offset form), out[16 ns].
__global__ void __launch_bounds__(256) k_mix_diag_aux(MixArrays a, double *out) {
    __shared__ double lds[10 * 256];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int is = blockIdx.x;
    const long long b0 = a.off[is], b1 = a.off[is + 1];
    double v[2] = {0, 0};
    double m[10] = {1e5, -1e5, 1e5, -1e5, 1e5, -1e5, 1e5, -1e5, 1e7, 1};
    for (long long i = b0 + t; i < b1; i += 256) {
        const double xi = a.x[i], yi = a.y[i], pxi = a.px[i], pyi = a.py[i];
        const double gi = a.gref + (double)a.dg[i];
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
    bsum<256, 2>(v, lds);
#pragma unroll
    for (int k = 0; k < 10; k++)
        for (int o = 32; o > 0; o >>= 1) {
            const double u = __shfl_xor(m[k], o);
            m[k] = (k & 1) ? fmax(m[k], u) : fmin(m[k], u);
        }
    __syncthreads();
    if (lane == 0)
        for (int k = 0; k < 10; k++) lds[wv * 10 + k] = m[k];
    __syncthreads();
    if (t == 0) {
        double *o = out + 16 * (long long)is;
        o[0] = v[0];
        o[1] = v[1];
        for (int k = 0; k < 10; k++) {
            double r = lds[k];
            for (int q = 1; q < 4; q++) r = (k & 1) ? fmax(r, lds[q * 10 + k]) : fmin(r, lds[q * 10 + k]);
            o[2 + k] = r;
        }
        const int ng = a.ng;
        const float2 c = a.field[a.st->cur][(long long)((is + a.st->first) % a.ns) * ng * ng + (ng / 2) * ng + ng / 2];
        o[12] = c.x;
        o[13] = c.y;
    }
}

void launch_mix_diag_aux(hipStream_t s, const MixArrays &a, double *out) {
    hipLaunchKernelGGL(k_mix_diag_aux, dim3(a.ns), dim3(256), 0, s, a, out);
}

void launch_mix_diag(hipStream_t s, const MixArrays &a, double *beam8, double *field8) {
    hipLaunchKernelGGL(k_mix_diag_beam, dim3(a.ns), dim3(256), 0, s, a, beam8);
    hipLaunchKernelGGL(k_mix_diag_field, dim3(a.ns), dim3(256), 0, s, a, field8, 0);
}

void launch_mix_sort_index(hipStream_t s, const MixArrays &a, double invslen, int *atar,
                           unsigned long long *moved) {
    hipLaunchKernelGGL(k_mix_sort_index, dim3(a.ns), dim3(256), 0, s, a, invslen, atar, moved);
}

}  // namespace g4mi
