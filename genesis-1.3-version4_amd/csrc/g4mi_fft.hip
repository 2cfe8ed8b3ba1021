// Complex FP64 GEMMs of the spectral field solver (FieldSolverFFT,
// src/Core/FieldSolverFFT.cpp:62-119), on v_mfma_f64_16x16x4_f64.
//
// One workgroup computes a T x T output tile (T = 16 NW, NW wavefronts of
// 16 rows each; every wavefront covers all NB = NW column blocks of 16).
// k runs in chunks of 16 staged through LDS as complex values; the next
// chunk's global loads are issued before the current chunk's MFMAs.  A
// complex product is four real MFMAs into separate real and imaginary
// accumulators (re += ar br - ai bi, im += ar bi + ai br).
//
// MFMA operand layout (16x16x4, FP64, measured on gfx950): lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator element v of lane l
// is D[4 v + l / 16][l % 16].
//
// Blocks are numbered XCD-aware: workgroups go to the 8 XCDs round-robin
// (block b on XCD b % 8), so the blocks that share a data tile (all column
// tiles of a row tile; all tiles of a slice) are given ids that are equal
// mod 8 and read that tile through one L2.
#include <algorithm>

#include "g4mi_fft.h"

namespace g4mi {

typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bool in_box(const int4 &bx, int x, int y) {
    return x >= bx.x && x <= bx.z && y >= bx.y && y <= bx.w;
}

// GAUSS: three real products per complex one, P1 = ar br, P2 = ai bi,
// P3 = (ar + ai)(br + bi), re = P1 - P2, im = P3 - P1 - P2 (a quarter fewer
// MFMAs, one more accumulator); otherwise four.  KC: k per LDS chunk.
template <int NW, bool ROWS, bool GAUSS, int KC>
__global__ void __launch_bounds__(64 * NW) k_cgemm(CgemmArgs g, int ntile_c, int ntile_r) {
    constexpr int T = 16 * NW, NT = 64 * NW, NB = NW;
    constexpr int PA = T + 1;                 // LDS pitch (complex) of a k row
    constexpr int LPT = (T * KC) / NT;        // loads per thread per operand (KC / 4)
    static_assert((T * KC) % NT == 0, "tile/thread mismatch");
    __shared__ double2 As[KC * PA];
    __shared__ double2 Bs[KC * PA];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int ng = g.ng;
    const long long plane = (long long)ng * ng;

    // logical tile of this block
    int s = 0, r0 = 0, c0 = 0;
    long long rr0 = 0;
    {
        const int b = blockIdx.x;
        const int per = 8 * ntile_c * (ROWS ? 1 : ntile_r);
        const int grp = b / per, rem = b % per;
        const int major = grp * 8 + (rem & 7), tile = rem >> 3;
        if (ROWS) {
            rr0 = (long long)major * T;
            if (rr0 >= g.nrows) return;  // uniform per block: padding of the last group
            c0 = tile * T;
        } else {
            s = major;
            if (s >= g.ns) return;
            r0 = (tile % ntile_r) * T;
            c0 = (tile / ntile_r) * T;
        }
    }
    const double2 *bsl = ROWS ? nullptr : g.b + (long long)s * plane;

    double2 ra[LPT], rb[LPT];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < LPT; q++) {
            const int e = t + q * NT;
            const int row = e / KC, kk = e % KC, l = k0 + kk;
            double2 va = make_double2(0.0, 0.0), vb = make_double2(0.0, 0.0);
            if (ROWS) {
                const long long r = rr0 + row;
                if (r < g.nrows && l < ng) {
                    bool ok = true;
                    if (g.mask_a) ok = in_box(g.box[r / ng], l, (int)(r % ng));
                    if (ok) va = g.a[r * ng + l];
                }
                const int j = c0 + row;  // B(k = l, n = j) = mat[j][l]
                if (j < ng && l < ng) vb = g.mat[(long long)j * ng + l];
            } else {
                const int j = r0 + row;  // A(m = j, k = l) = mat[j][l]
                if (j < ng && l < ng) va = g.mat[(long long)j * ng + l];
                const int kb = e / T, col = e % T, lb = k0 + kb, x = c0 + col;  // B(k, n) = b[s][l][x]
                if (lb < ng && x < ng) vb = bsl[(long long)lb * ng + x];
            }
            ra[q] = va;
            rb[q] = vb;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < LPT; q++) {
            const int e = t + q * NT;
            As[(e % KC) * PA + e / KC] = ra[q];
            if (ROWS)
                Bs[(e % KC) * PA + e / KC] = rb[q];
            else
                Bs[(e / T) * PA + e % T] = rb[q];
        }
    };

    f64x4 cr[NB], ci[NB], c3[GAUSS ? NB : 1];
#pragma unroll
    for (int i = 0; i < NB; i++) {
        cr[i] = f64x4{0.0, 0.0, 0.0, 0.0};
        ci[i] = f64x4{0.0, 0.0, 0.0, 0.0};
        if (GAUSS) c3[i] = f64x4{0.0, 0.0, 0.0, 0.0};
    }
    const int nk = (ng + KC - 1) / KC;
    fetch(0);
    for (int kc = 0; kc < nk; kc++) {
        __syncthreads();  // previous chunk's reads done
        stash();
        __syncthreads();
        if (kc + 1 < nk) fetch((kc + 1) * KC);
#pragma unroll
        for (int ks = 0; ks < KC / 4; ks++) {
            const int kk = ks * 4 + (lane >> 4);
            const double2 av = As[kk * PA + w * 16 + (lane & 15)];
            if (GAUSS) {
                const double asum = av.x + av.y;
#pragma unroll
                for (int i = 0; i < NB; i++) {
                    const double2 bv = Bs[kk * PA + i * 16 + (lane & 15)];
                    cr[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.x, cr[i], 0, 0, 0);
                    ci[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv.y, ci[i], 0, 0, 0);
                    c3[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(asum, bv.x + bv.y, c3[i], 0, 0, 0);
                }
            } else {
                const double nai = -av.y;
#pragma unroll
                for (int i = 0; i < NB; i++) {
                    const double2 bv = Bs[kk * PA + i * 16 + (lane & 15)];
                    cr[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.x, cr[i], 0, 0, 0);
                    cr[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, bv.y, cr[i], 0, 0, 0);
                    ci[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.y, ci[i], 0, 0, 0);
                    ci[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv.x, ci[i], 0, 0, 0);
                }
            }
        }
    }

    // epilogue
    const int ii = ROWS ? 0 : (g.ring ? (s - g.first + g.ns) % g.ns : s);
    int4 bx = make_int4(1, 1, 0, 0);
    if (!ROWS && g.src) bx = g.box[ii];
#pragma unroll
    for (int i = 0; i < NB; i++) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = w * 16 + 4 * v + (lane >> 4), col = i * 16 + (lane & 15);
            double re, im;
            if (GAUSS) {
                re = (cr[i][v] - ci[i][v]) * g.scale;
                im = (c3[i][v] - cr[i][v] - ci[i][v]) * g.scale;
            } else {
                re = cr[i][v] * g.scale;
                im = ci[i][v] * g.scale;
            }
            if (ROWS) {
                const long long r = rr0 + row;
                const int j = c0 + col;
                if (r < g.nrows && j < ng) g.c[r * ng + j] = make_double2(re, im);
            } else {
                const int j = r0 + row, x = c0 + col;
                if (j < ng && x < ng) {
                    const long long o = (long long)j * ng + x;
                    if (g.ptw) {
                        const double f = g.ptw[o];
                        re *= f;
                        im *= f;
                    }
                    if (g.add) {
                        const double2 a2 = g.add[(long long)ii * plane + o];
                        re += 2. * a2.x;
                        im += 2. * a2.y;
                    }
                    if (g.src && in_box(bx, x, j)) {
                        const double2 s2 = g.src[(long long)ii * plane + o];
                        re += 2. * s2.x;
                        im += 2. * s2.y;
                    }
                    g.c[(long long)s * plane + o] = make_double2(re, im);
                }
            }
        }
    }
}

int cgemm_tile(int ng) {
    int best = 80, waste = 1 << 30;
    for (int T : {80, 64, 48, 32}) {
        const int pad = ((ng + T - 1) / T) * T;
        if (pad < waste) {
            waste = pad;
            best = T;
        }
    }
    return best;
}

template <bool ROWS, bool GAUSS, int KC>
static void launch_variant(hipStream_t st, const CgemmArgs &g, int T, dim3 grid, int ntc, int ntr) {
    switch (T) {
        case 80: hipLaunchKernelGGL((k_cgemm<5, ROWS, GAUSS, KC>), grid, dim3(320), 0, st, g, ntc, ntr); break;
        case 64: hipLaunchKernelGGL((k_cgemm<4, ROWS, GAUSS, KC>), grid, dim3(256), 0, st, g, ntc, ntr); break;
        case 48: hipLaunchKernelGGL((k_cgemm<3, ROWS, GAUSS, KC>), grid, dim3(192), 0, st, g, ntc, ntr); break;
        default: hipLaunchKernelGGL((k_cgemm<2, ROWS, GAUSS, KC>), grid, dim3(128), 0, st, g, ntc, ntr); break;
    }
}

template <bool ROWS>
static int launch_cgemm(hipStream_t st, const CgemmArgs &g) {
    if (g.ng < 2 || !g.mat || !g.c || (ROWS ? !g.a : !g.b)) {
        set_error("cgemm: bad arguments");
        return G4MI_E_ARG;
    }
    const int T = cgemm_tile(g.ng);
    const int ntc = (g.ng + T - 1) / T;
    const int ntr = ROWS ? 1 : ntc;
    const long long major = ROWS ? (g.nrows + T - 1) / T : g.ns;
    const long long blocks = ((major + 7) / 8) * 8 * ntc * ntr;
    if (blocks > 0x7fffffffLL) {
        set_error("cgemm: problem too large for one launch");
        return G4MI_E_ARG;
    }
    const dim3 grid((unsigned)blocks);
    switch (g.variant) {
        case 1: launch_variant<ROWS, true, 16>(st, g, T, grid, ntc, ntr); break;
        case 2: launch_variant<ROWS, false, 32>(st, g, T, grid, ntc, ntr); break;
        case 3: launch_variant<ROWS, true, 32>(st, g, T, grid, ntc, ntr); break;
        default: launch_variant<ROWS, false, 16>(st, g, T, grid, ntc, ntr); break;
    }
    G4_HIP(hipGetLastError());
    return G4MI_OK;
}

// Partial moments of |F|^2 over the FFT index: block (p, s) of a slice's
// np blocks covers cells [p P, (p + 1) P), P = ceil(ng^2 / np), so that a
// window of few slices still spreads over many workgroups.
__global__ void __launch_bounds__(256) k_fft_moments(const double2 *F, int ng, int np, double *part) {
    __shared__ double red[5][256];
    const int s = blockIdx.y, p = blockIdx.x, t = threadIdx.x;
    const double2 *f = F + (long long)s * ng * ng;
    const double shift = -0.5 * (double)(ng - 1);
    const int nn = ng * ng, chunk = (nn + np - 1) / np;
    const int k0 = p * chunk, k1 = min(nn, k0 + chunk);
    double acc[5] = {0, 0, 0, 0, 0};
    for (int k = k0 + t; k < k1; k += 256) {
        const int ky = k / ng, kx = k % ng;
        // grid point of FFT index k: (k - (n+1)/2) mod n
        const double dx = (double)((kx - (ng + 1) / 2 + ng) % ng) + shift;
        const double dy = (double)((ky - (ng + 1) / 2 + ng) % ng) + shift;
        const double2 v = f[k];
        const double w = v.x * v.x + v.y * v.y;
        acc[0] += w;
        acc[1] += dx * w;
        acc[2] += dx * dx * w;
        acc[3] += dy * w;
        acc[4] += dy * dy * w;
    }
    for (int q = 0; q < 5; q++) red[q][t] = acc[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
            for (int q = 0; q < 5; q++) red[q][t] += red[q][t + h];
        __syncthreads();
    }
    if (t < 5) part[((long long)s * np + p) * 5 + t] = red[t][0];
}

// out[5 s + q] = the slice's partial sums in block order
__global__ void k_fft_moments_fold(const double *part, int ns, int np, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 5 * ns) return;
    const int s = i / 5, q = i % 5;
    double v = 0;
    for (int p = 0; p < np; p++) v += part[((long long)s * np + p) * 5 + q];
    out[i] = v;
}

int fft_moments_parts(int ns) { return std::max(1, std::min(kFftMomentParts, (512 + ns - 1) / ns)); }

void launch_fft_moments(hipStream_t s, int ns, const double2 *F, int ng, double *part, double *out) {
    if (ns <= 0) return;
    const int np = fft_moments_parts(ns);
    hipLaunchKernelGGL(k_fft_moments, dim3(np, ns), dim3(256), 0, s, F, ng, np, part);
    hipLaunchKernelGGL(k_fft_moments_fold, dim3((5 * ns + 255) / 256), dim3(256), 0, s, part, ns, np, out);
}

__global__ void k_widen_field(const float2 *__restrict__ in, double2 *__restrict__ out, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = make_double2((double)in[i].x, (double)in[i].y);
}

void launch_widen_field(hipStream_t s, const float2 *in, double2 *out, long long n) {
    if (n <= 0) return;
    const long long blocks = std::min<long long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_widen_field, dim3((unsigned)blocks), dim3(256), 0, s, in, out, n);
}

int launch_cgemm_rows(hipStream_t s, const CgemmArgs &g) { return launch_cgemm<true>(s, g); }
int launch_cgemm_cols(hipStream_t s, const CgemmArgs &g) { return launch_cgemm<false>(s, g); }

// ---- propagation by FFT convolution ---------------------------------------
//
// Every line (a row, or a column of one slice) of n values is convolved
// circularly with the propagator kernel m: zero-padded to L >= 2n - 1, the
// cyclic convolution of length L with the kernel placed at d mod L for
// |d| < n equals the circular one of length n on the first n outputs.  The
// FFTs are Stockham autosort passes (Govindaraju et al., "High Performance
// Discrete Fourier Transforms on Graphics Processors", SC08): L = R^NST with
// R = L / 64, so each of a wavefront's 64 lanes does one radix-R butterfly
// per pass on R values held in registers, reading them at stride 64 and
// writing them at stride ns in natural order for the next pass.
//
// The last forward pass leaves lane j holding X[j + 64 r], which is what the
// first inverse pass reads: the product with the kernel's transform (hk,
// already divided by L) happens in registers, with no exchange.  As n <= L/2,
// the first pass's inputs r >= R/2 are zero and the last pass's outputs
// r >= R/2 are not needed.  Rows are read and written by the lanes directly
// (64 consecutive values per access); columns go through a tile of 8
// adjacent columns in LDS so global accesses stay row-contiguous, with the
// column groups of a slice numbered onto one XCD so the 128-byte lines they
// share are fetched into one L2.  Within an LDS line, entry p is stored at
// p ^ ((p >> 3) & 31): with it the stride-ns writes and stride-64 reads of
// the radix-8 exchanges touch every bank once per pass whether the LDS
// serves 8, 16 or 32 lanes of 16 bytes per cycle.

__device__ __forceinline__ double2 conv_cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 conv_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 conv_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// -i z (forward) or +i z (inverse)
template <bool INV>
__device__ __forceinline__ double2 conv_rot(double2 z) {
    return INV ? make_double2(-z.y, z.x) : make_double2(z.y, -z.x);
}
__device__ __forceinline__ int conv_swz(int p) { return p ^ ((p >> 3) & 31); }

// in-register DFT of length R (2, 4 or 8), sign - (forward) or + (INV)
template <int R, bool INV>
__device__ __forceinline__ void conv_dft(double2 (&v)[R]) {
    if constexpr (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = conv_add(a, b);
        v[1] = conv_sub(a, b);
    } else if constexpr (R == 4) {
        const double2 t0 = conv_add(v[0], v[2]), t1 = conv_sub(v[0], v[2]);
        const double2 t2 = conv_add(v[1], v[3]), t3 = conv_rot<INV>(conv_sub(v[1], v[3]));
        v[0] = conv_add(t0, t2);
        v[2] = conv_sub(t0, t2);
        v[1] = conv_add(t1, t3);
        v[3] = conv_sub(t1, t3);
    } else {
        double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        conv_dft<4, INV>(e);
        conv_dft<4, INV>(o);
        const double c = 0.70710678118654752440;  // sqrt(1/2)
        // o[k] *= exp(-+ 2 pi i k / 8)
        o[1] = INV ? make_double2(c * (o[1].x - o[1].y), c * (o[1].x + o[1].y))
                   : make_double2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));
        o[2] = conv_rot<INV>(o[2]);
        o[3] = INV ? make_double2(-c * (o[3].x + o[3].y), c * (o[3].x - o[3].y))
                   : make_double2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = conv_add(e[k], o[k]);
            v[k + 4] = conv_sub(e[k], o[k]);
        }
    }
}

// one Stockham pass of span ns on lane j's registers: twiddles
// w^(r (j mod ns) L / (ns R)), w = exp(-+ 2 pi i / L), from tw[] for r = 1
// and by products for higher r; then the radix-R butterfly
template <int R, bool INV>
__device__ __forceinline__ void conv_pass(double2 (&v)[R], int ns, int j, const double2 *__restrict__ tw) {
    constexpr int L = 64 * R;
    if (ns > 1) {
        double2 w1 = tw[(j % ns) * (L / (ns * R))];
        if (INV) w1.y = -w1.y;
        double2 wr = w1;
        v[1] = conv_cmul(v[1], wr);
#pragma unroll
        for (int r = 2; r < R; r++) {
            wr = conv_cmul(wr, w1);
            v[r] = conv_cmul(v[r], wr);
        }
    }
    conv_dft<R, INV>(v);
}

// workgroup barrier that orders LDS only: global loads issued before it (the
// next line's prefetch) stay in flight
__device__ __forceinline__ void conv_lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// a whole transform on registers, exchanging through this wavefront's LDS
// line d between passes; in: v[r] = x[j + 64 r]; out: v[r] = X[j + 64 r]
template <int R, int NST, bool INV>
__device__ __forceinline__ void conv_fft(double2 (&v)[R], double2 *d, int j, const double2 *__restrict__ tw) {
    int ns = 1;
#pragma unroll
    for (int st = 0; st < NST; st++) {
        conv_pass<R, INV>(v, ns, j, tw);
        if (st < NST - 1) {
            const int dst = (j / ns) * ns * R + j % ns;
#pragma unroll
            for (int r = 0; r < R; r++) d[conv_swz(dst + r * ns)] = v[r];
            conv_lds_barrier();
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = d[conv_swz(j + 64 * r)];
            conv_lds_barrier();
        }
        ns *= R;
    }
}

__device__ __forceinline__ bool conv_in_box(const int4 &bx, int x, int y) {
    return x >= bx.x && x <= bx.z && y >= bx.y && y <= bx.w;
}

template <bool COLS>
struct ConvLines {
    static constexpr int value = 4;  // lines (wavefronts) per workgroup
};

// One chunk of NL lines per workgroup (rows: NL consecutive rows; columns:
// NL adjacent columns of one slice).  Workgroups go to the 8 XCDs
// round-robin; the chunks are numbered so that XCD b % 8 takes a contiguous
// range of them, and the column groups of a slice, which share 128-byte
// lines, are read through one L2.
template <int R, int NST, bool COLS>
__global__ void __launch_bounds__(64 * ConvLines<COLS>::value) k_fft_conv(ConvArgs a, int nchunks) {
    constexpr int L = 64 * R, NL = ConvLines<COLS>::value, NT = 64 * NL, LP = L + 2, NP = R / 2;
    __shared__ double2 buf[NL * LP];
    const int t = threadIdx.x, j = t & 63, w = t >> 6;
    const int ng = a.ng;
    const long long nn = (long long)ng * ng;
    const int c = (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
    if (c >= nchunks) return;  // padding of the grid to a multiple of 8
    double2 *d = buf + w * LP;
    double2 v[R];
#pragma unroll
    for (int r = NP; r < R; r++) v[r] = make_double2(0., 0.);
    const int cpb = (ng + NL - 1) / NL;
    const int s = COLS ? c / cpb : 0, x0 = COLS ? (c % cpb) * NL : 0;
    const long long g = COLS ? 0 : (long long)c * NL + w;
    if (COLS) {
        // tile of NL adjacent columns, positions [0, L/2) (zero from ng on)
#pragma unroll
        for (int q = 0; q < NP; q++) {
            const int e = t + NT * q, line = e % NL, pos = e / NL, x = x0 + line;
            double2 u = make_double2(0., 0.);
            if (pos < ng && x < ng) u = a.in[(long long)s * nn + (long long)pos * ng + x];
            buf[line * LP + conv_swz(pos)] = u;
        }
        conv_lds_barrier();
#pragma unroll
        for (int r = 0; r < NP; r++) v[r] = d[conv_swz(j + 64 * r)];
        conv_lds_barrier();
    } else {
        int4 mb = make_int4(0, 0, ng - 1, ng - 1);
        if (a.mask && g < a.nrows) mb = a.box[g / ng];
        const int yrow = (int)(g % ng);
        const bool rin = yrow >= mb.y && yrow <= mb.w;
#pragma unroll
        for (int r = 0; r < NP; r++) {
            const int p = j + 64 * r;
            const bool ok = g < a.nrows && p < ng && rin && p >= mb.x && p <= mb.z;
            v[r] = ok ? a.in[g * ng + p] : make_double2(0., 0.);
        }
    }
    if (a.pre) {
#pragma unroll
        for (int r = 0; r < NP; r++)
            if (j + 64 * r < ng) v[r] = conv_cmul(v[r], a.pre[j + 64 * r]);
    }
    conv_fft<R, NST, false>(v, d, j, a.tw);
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = conv_cmul(v[r], a.hk[j + 64 * r]);
    conv_fft<R, NST, true>(v, d, j, a.tw);
    if (a.post) {
#pragma unroll
        for (int r = 0; r < NP; r++)
            if (j + 64 * r < ng) v[r] = conv_cmul(v[r], a.post[j + 64 * r]);
    }
    if (!COLS) {
        if (g < a.nrows) {
#pragma unroll
            for (int r = 0; r < NP; r++) {
                const int p = j + 64 * r;
                if (p < ng) a.out[g * ng + p] = v[r];
            }
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < NP; r++) d[conv_swz(j + 64 * r)] = v[r];
    conv_lds_barrier();
    const int ii = a.ring ? (s - a.first + a.ns) % a.ns : s;
    int4 bx = make_int4(1, 1, 0, 0);
    if (a.src) bx = a.box[ii];
    for (int e = t; e < NL * ng; e += NT) {
        const int line = e % NL, y = e / NL, x = x0 + line;
        if (x >= ng) continue;
        double2 u = buf[line * LP + conv_swz(y)];
        const long long o = (long long)y * ng + x;
        if (a.scale != 1.0) {
            u.x *= a.scale;
            u.y *= a.scale;
        }
        if (a.ptw) {
            const double f = a.ptw[o];
            u.x *= f;
            u.y *= f;
        }
        if (a.add) {
            const double2 a2 = a.add[(long long)ii * nn + o];
            u.x += 2. * a2.x;
            u.y += 2. * a2.y;
        }
        if (a.src && conv_in_box(bx, x, y)) {
            const double2 s2 = a.src[(long long)ii * nn + o];
            u.x += 2. * s2.x;
            u.y += 2. * s2.y;
        }
        a.out[(long long)s * nn + o] = u;
    }
}

template <bool COLS>
static int launch_fft_conv(hipStream_t st, const ConvArgs &a) {
    if (a.ng < 2 || !a.in || !a.out || !a.hk || !a.tw || a.in == a.out || a.L < 2 * a.ng - 1 ||
        (a.L != 128 && a.L != 256 && a.L != 512)) {
        set_error("fft_conv: bad arguments");
        return G4MI_E_ARG;
    }
    constexpr int NL = ConvLines<COLS>::value;
    const long long chunks = COLS ? (long long)a.ns * ((a.ng + NL - 1) / NL) : (a.nrows + NL - 1) / NL;
    if (chunks < 1) return G4MI_OK;
    if (chunks > 0x7ffffff0LL) {
        set_error("fft_conv: problem too large for one launch");
        return G4MI_E_ARG;
    }
    const long long blocks = (chunks + 7) / 8 * 8;  // one chunk per workgroup, every XCD the same count
    const dim3 grid((unsigned)blocks), block(64 * NL);
    const int nc = (int)chunks;
    switch (a.L) {
        case 128: hipLaunchKernelGGL((k_fft_conv<2, 7, COLS>), grid, block, 0, st, a, nc); break;
        case 256: hipLaunchKernelGGL((k_fft_conv<4, 4, COLS>), grid, block, 0, st, a, nc); break;
        default: hipLaunchKernelGGL((k_fft_conv<8, 3, COLS>), grid, block, 0, st, a, nc); break;
    }
    G4_HIP(hipGetLastError());
    return G4MI_OK;
}

int launch_fft_conv_rows(hipStream_t s, const ConvArgs &a) { return launch_fft_conv<false>(s, a); }
int launch_fft_conv_cols(hipStream_t s, const ConvArgs &a) { return launch_fft_conv<true>(s, a); }

}  // namespace g4mi
