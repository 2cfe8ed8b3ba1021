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

__global__ void __launch_bounds__(256) k_fft_moments(const double2 *__restrict__ F, int ng,
                                                     double *__restrict__ out) {
    __shared__ double red[5][256];
    const int s = blockIdx.x, t = threadIdx.x;
    const double2 *f = F + (long long)s * ng * ng;
    const double shift = -0.5 * (double)(ng - 1);
    double acc[5] = {0, 0, 0, 0, 0};
    for (int k = t; k < ng * ng; k += 256) {
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
    if (t < 5) out[5 * s + t] = red[t][0];
}

void launch_fft_moments(hipStream_t s, int ns, const double2 *F, int ng, double *out) {
    if (ns > 0) hipLaunchKernelGGL(k_fft_moments, dim3(ns), dim3(256), 0, s, F, ng, out);
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

}  // namespace g4mi
