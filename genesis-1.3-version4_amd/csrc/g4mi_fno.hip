// FNO surrogate of the per-slice field step: BF16 MFMA GEMM, the network's
// forward and backward passes, Adam, the training set captured from ADI
// steps, and batched inference over the slices of a context.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "g4mi_fno.h"
#include "g4mi_fno_api.h"

namespace g4mi {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// ---------------------------------------------------------------------------
// elementwise helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float gelu(float x) {
    const float c = 0.7978845608028654f;  // sqrt(2/pi)
    return 0.5f * x * (1.f + tanhf(c * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float gelu_d(float x) {
    const float c = 0.7978845608028654f;
    const float u = c * (x + 0.044715f * x * x * x);
    const float th = tanhf(u);
    return 0.5f * (1.f + th) + 0.5f * x * (1.f - th * th) * c * (1.f + 3.f * 0.044715f * x * x);
}

// ---------------------------------------------------------------------------
// GEMM: (32 WM) x (32 WN) output tile per workgroup of WM x WN wavefronts,
// each wavefront one 32x32 block; k in chunks of 32 staged through LDS as
// BF16 (global reads as float4 where the contiguous dimension allows),
// gfx950's v_mfma_f32_32x32x16_bf16 (two per chunk) with FP32 accumulation.
// MFMA operand layout (32x32x16): lane l holds A row l%32 and B column l%32
// at k = 8 (l/32) .. 8 (l/32) + 7 (one 16-byte LDS read each); accumulator
// register v of lane l is C row 8 (v/4) + 4 (l/32) + v%4, column l%32.
// ---------------------------------------------------------------------------
constexpr int GK = 32;
constexpr int GP = GK + 8;  // LDS row pitch (bf16): 16-byte aligned rows

__device__ __forceinline__ void st_bf4(unsigned short *p, float4 v) {
    // 4 consecutive bf16 in one 8-byte LDS store
    const unsigned lo = (unsigned)f2bf(v.x) | ((unsigned)f2bf(v.y) << 16);
    const unsigned hi = (unsigned)f2bf(v.z) | ((unsigned)f2bf(v.w) << 16);
    *(uint2 *)p = make_uint2(lo, hi);
}

// four consecutive floats p[0..3] of which the first `n` (0..4) exist
__device__ __forceinline__ float4 ld4(const float *p, int n, bool vec) {
    if (vec && n >= 4) return *(const float4 *)p;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (n > 0) v.x = p[0];
    if (n > 1) v.y = p[1];
    if (n > 2) v.z = p[2];
    if (n > 3) v.w = p[3];
    return v;
}

template <int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN) k_gemm_bf16(GemmArgs g, int vecA, int vecB) {
    constexpr int NT = 64 * WM * WN, TM = 32 * WM, TN = 32 * WN;
    __shared__ __align__(16) unsigned short As[TM * GP];   // [m][k]
    __shared__ __align__(16) unsigned short Bs[TN * GP];   // [n][k]
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wm = wv / WN, wn = wv % WN;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const long long z = blockIdx.z;
    const float *A = g.A + z * g.sA, *B = g.B + z * g.sB;
    float *C = g.C + z * g.sC;
    int K = g.K;
    if (g.ktot > 0) K = (int)std::min((long long)g.K, g.ktot - z * g.K);
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; v++) acc[v] = 0.f;
    for (int k0 = 0; k0 < K; k0 += GK) {
        // A tile: TM x GK in units of 4 along the contiguous dimension
        for (int e = t; e < TM * GK / 4; e += NT) {
            if (!g.transA) {  // A[m][k]: contiguous along k
                const int mm = e / (GK / 4), kk = (e % (GK / 4)) * 4;
                const int m = m0 + mm, k = k0 + kk;
                const float4 v = m < g.M ? ld4(A + (long long)m * g.lda + k, K - k, vecA)
                                         : make_float4(0.f, 0.f, 0.f, 0.f);
                st_bf4(&As[mm * GP + kk], v);
            } else {  // A[k][m]: contiguous along m
                const int kk = e / (TM / 4), mm = (e % (TM / 4)) * 4;
                const int m = m0 + mm, k = k0 + kk;
                const float4 v = k < K ? ld4(A + (long long)k * g.lda + m, g.M - m, vecA)
                                       : make_float4(0.f, 0.f, 0.f, 0.f);
                As[(mm + 0) * GP + kk] = f2bf(v.x);
                As[(mm + 1) * GP + kk] = f2bf(v.y);
                As[(mm + 2) * GP + kk] = f2bf(v.z);
                As[(mm + 3) * GP + kk] = f2bf(v.w);
            }
        }
        for (int e = t; e < TN * GK / 4; e += NT) {
            if (g.transB) {  // B[n][k]: contiguous along k
                const int nn = e / (GK / 4), kk = (e % (GK / 4)) * 4;
                const int n = n0 + nn, k = k0 + kk;
                const float4 v = n < g.N ? ld4(B + (long long)n * g.ldb + k, K - k, vecB)
                                         : make_float4(0.f, 0.f, 0.f, 0.f);
                st_bf4(&Bs[nn * GP + kk], v);
            } else {  // B[k][n]: contiguous along n
                const int kk = e / (TN / 4), nn = (e % (TN / 4)) * 4;
                const int n = n0 + nn, k = k0 + kk;
                const float4 v = k < K ? ld4(B + (long long)k * g.ldb + n, g.N - n, vecB)
                                       : make_float4(0.f, 0.f, 0.f, 0.f);
                Bs[(nn + 0) * GP + kk] = f2bf(v.x);
                Bs[(nn + 1) * GP + kk] = f2bf(v.y);
                Bs[(nn + 2) * GP + kk] = f2bf(v.z);
                Bs[(nn + 3) * GP + kk] = f2bf(v.w);
            }
        }
        __syncthreads();
        const int mr = wm * 32 + (lane & 31), nc = wn * 32 + (lane & 31);
#pragma unroll
        for (int kk = 0; kk < GK; kk += 16) {
            const int kq = kk + 8 * (lane >> 5);
            const bf16x8 a = *(const bf16x8 *)&As[mr * GP + kq];
            const bf16x8 b = *(const bf16x8 *)&Bs[nc * GP + kq];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int col = n0 + wn * 32 + (lane & 31);
    if (col >= g.N) return;
    const float bcol = g.bias ? g.bias[col] : 0.f;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int row = m0 + wm * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (row < g.M) {
            const long long o = (long long)row * g.ldc + col;
            float r = g.alpha * acc[v];
            if (g.beta != 0.f) r += g.beta * C[o];
            r += bcol;
            if (g.dmul) r *= gelu_d(g.dmul[z * g.sC + o]);
            C[o] = r;
            if (g.act_out) g.act_out[z * g.sC + o] = g.act ? gelu(r) : r;
        }
    }
}

// column sums of d [rows][ch] (ch <= 256): stage 1, block b sums rows
// b, b + nblk, ... into part[b][ch] (coalesced: lane -> (row, channel));
// stage 2 adds the nblk partials in block order
constexpr int kColBlocks = 256;
__global__ void __launch_bounds__(256) k_colsum_part(const float *d, long long rows, int ch, double *part) {
    __shared__ double red[256];
    const int t = threadIdx.x;
    const int rpb = 256 / ch;  // rows per pass of the block
    double s = 0;
    if (t < rpb * ch) {
        const int c = t % ch;
        for (long long r = (long long)blockIdx.x * rpb + t / ch; r < rows; r += (long long)kColBlocks * rpb)
            s += d[r * ch + c];
    }
    red[t] = s;
    __syncthreads();
    if (t < ch) {
        double acc = 0;
        for (int q = 0; q < rpb; q++) acc += red[q * ch + t];
        part[(long long)blockIdx.x * ch + t] = acc;
    }
}
__global__ void k_colsum_fin(const double *part, int ch, float *out) {
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= ch) return;
    double s = 0;
    for (int b = 0; b < kColBlocks; b++) s += part[(long long)b * ch + c];
    out[c] = (float)s;
}

// sum of nz partial blocks of n floats: blockIdx.y = g adds blocks
// [g zc, min((g + 1) zc, nz)) into out[g][n]
__global__ void k_reduce_parts(const float *parts, int nz, int zc, long long n, float *out) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int z0 = blockIdx.y * zc, z1 = min(z0 + zc, nz);
    float s = 0;
    for (int z = z0; z < z1; z++) s += parts[(long long)z * n + i];
    out[blockIdx.y * n + i] = s;
}

// Ahat[b][ky][kx][2C] from T[b][2nm][2nm C] (y transform of the x transform):
// Re = Cos[ky][kx] - MSin[ky][nm + kx], Im = Cos[ky][nm + kx] + MSin[ky][kx]
__global__ void k_comb_fwd(const float *T, float *Ah, int B, int nm, int C) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long tot = (long long)B * nm * nm * C;
    if (i >= tot) return;
    const int c = (int)(i % C);
    const int kx = (int)((i / C) % nm);
    const int ky = (int)((i / ((long long)C * nm)) % nm);
    const int b = (int)(i / ((long long)C * nm * nm));
    const long long rowlen = 2LL * nm * C;
    const float *Tb = T + (long long)b * 2 * nm * rowlen;
    const float cr = Tb[(long long)ky * rowlen + (long long)kx * C + c];
    const float ci = Tb[(long long)ky * rowlen + (long long)(nm + kx) * C + c];
    const float sr = Tb[(long long)(nm + ky) * rowlen + (long long)kx * C + c];
    const float si = Tb[(long long)(nm + ky) * rowlen + (long long)(nm + kx) * C + c];
    float *o = Ah + (((long long)b * nm + ky) * nm + kx) * 2 * C;
    o[c] = cr - si;
    o[C + c] = ci + sr;
}

// adjoint of k_comb_fwd: dT from dAhat (every dT entry gets one term)
__global__ void k_comb_fwd_back(const float *dAh, float *dT, int B, int nm, int C) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long tot = (long long)B * nm * nm * C;
    if (i >= tot) return;
    const int c = (int)(i % C);
    const int kx = (int)((i / C) % nm);
    const int ky = (int)((i / ((long long)C * nm)) % nm);
    const int b = (int)(i / ((long long)C * nm * nm));
    const long long rowlen = 2LL * nm * C;
    float *Tb = dT + (long long)b * 2 * nm * rowlen;
    const float *d = dAh + (((long long)b * nm + ky) * nm + kx) * 2 * C;
    const float dr = d[c], di = d[C + c];
    Tb[(long long)ky * rowlen + (long long)kx * C + c] = dr;
    Tb[(long long)(nm + ky) * rowlen + (long long)(nm + kx) * C + c] = -dr;
    Tb[(long long)ky * rowlen + (long long)(nm + kx) * C + c] = di;
    Tb[(long long)(nm + ky) * rowlen + (long long)kx * C + c] = di;
}

// Y1[b][y][part][kx][C] from Z[b][2N][nm 2C]: Re = Zc[Br] - Zs[Bi], Im = Zc[Bi] + Zs[Br]
__global__ void k_comb_inv(const float *Z, float *Y1, int B, int N, int nm, int C) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long tot = (long long)B * N * nm * C;
    if (i >= tot) return;
    const int c = (int)(i % C);
    const int kx = (int)((i / C) % nm);
    const int y = (int)((i / ((long long)C * nm)) % N);
    const int b = (int)(i / ((long long)C * nm * N));
    const long long rl = 2LL * nm * C;
    const float *Zb = Z + (long long)b * 2 * N * rl;
    const float zcr = Zb[(long long)y * rl + (long long)kx * 2 * C + c];
    const float zci = Zb[(long long)y * rl + (long long)kx * 2 * C + C + c];
    const float zsr = Zb[(long long)(N + y) * rl + (long long)kx * 2 * C + c];
    const float zsi = Zb[(long long)(N + y) * rl + (long long)kx * 2 * C + C + c];
    float *o = Y1 + ((long long)b * N + y) * 2 * nm * C;
    o[(long long)kx * C + c] = zcr - zsi;
    o[(long long)(nm + kx) * C + c] = zci + zsr;
}

__global__ void k_comb_inv_back(const float *dY1, float *dZ, int B, int N, int nm, int C) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long tot = (long long)B * N * nm * C;
    if (i >= tot) return;
    const int c = (int)(i % C);
    const int kx = (int)((i / C) % nm);
    const int y = (int)((i / ((long long)C * nm)) % N);
    const int b = (int)(i / ((long long)C * nm * N));
    const long long rl = 2LL * nm * C;
    float *Zb = dZ + (long long)b * 2 * N * rl;
    const float *d = dY1 + ((long long)b * N + y) * 2 * nm * C;
    const float dr = d[(long long)kx * C + c], di = d[(long long)(nm + kx) * C + c];
    Zb[(long long)y * rl + (long long)kx * 2 * C + c] = dr;
    Zb[(long long)y * rl + (long long)kx * 2 * C + C + c] = di;
    Zb[(long long)(N + y) * rl + (long long)kx * 2 * C + c] = di;
    Zb[(long long)(N + y) * rl + (long long)kx * 2 * C + C + c] = -dr;
}

// Wreal[mode] = [[Rr, Ri], [-Ri, Rr]] (2C x 2C) from Rr, Ri [mode][C][C]
__global__ void k_build_wreal(const float *Rr, const float *Ri, float *W, int nmode, int C) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long tot = (long long)nmode * C * C;
    if (i >= tot) return;
    const int o = (int)(i % C), c = (int)((i / C) % C);
    const long long m = i / ((long long)C * C);
    const float rr = Rr[i], ri = Ri[i];
    float *w = W + m * 4LL * C * C;
    w[(long long)c * 2 * C + o] = rr;
    w[(long long)c * 2 * C + C + o] = ri;
    w[(long long)(C + c) * 2 * C + o] = -ri;
    w[(long long)(C + c) * 2 * C + C + o] = rr;
}

// fold dWreal into dRr, dRi
__global__ void k_fold_wreal(const float *dW, float *dRr, float *dRi, int nmode, int C) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long tot = (long long)nmode * C * C;
    if (i >= tot) return;
    const int o = (int)(i % C), c = (int)((i / C) % C);
    const long long m = i / ((long long)C * C);
    const float *w = dW + m * 4LL * C * C;
    dRr[i] = w[(long long)c * 2 * C + o] + w[(long long)(C + c) * 2 * C + C + o];
    dRi[i] = w[(long long)c * 2 * C + C + o] - w[(long long)(C + c) * 2 * C + o];
}

// mean squared error of out [n] against tgt and its gradient 2 (out - tgt) / n
__global__ void __launch_bounds__(256) k_mse(const float *out, const float *tgt, float *dout, long long n,
                                             double *loss) {
    __shared__ double red[256];
    double s = 0;
    const float inv = 2.f / (float)n;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const float d = out[i] - tgt[i];
        s += (double)d * d;
        dout[i] = inv * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / (double)n);
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Batch assembly from the pair set: in[b][p][4] = (E, S) / scale_b,
// tgt[b][p][2] = (E_next - E - 2 S) / scale_b, scale_b = sqrt(<|E|^2 + |S|^2>)
// over the slice.  The ADI step is E_next = E + 2 S + (diffraction of E and
// S, first order in the step's i dz / (4 ks dgrid^2)): the network learns
// the diffraction only, the source's 2 S passes through exactly.
// pairs: [n][P][6] = Er Ei Sr Si Nr Ni.
__global__ void __launch_bounds__(256) k_scale(const float *pairs, const int *idx, int P, float *scale) {
    __shared__ double red[256];
    const int b = blockIdx.x;
    const float *pr = pairs + (long long)idx[b] * P * 6;
    double s = 0;
    for (int p = threadIdx.x; p < P; p += 256) {
        const float *q = pr + (long long)p * 6;
        s += (double)q[0] * q[0] + (double)q[1] * q[1] + (double)q[2] * q[2] + (double)q[3] * q[3];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) scale[b] = (float)(sqrt(red[0] / P) + 1e-30);
}

__global__ void k_assemble(const float *pairs, const int *idx, const float *scale, int B, int P,
                           float *in, float *tgt) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)B * P) return;
    const int b = (int)(i / P), p = (int)(i % P);
    const float *q = pairs + ((long long)idx[b] * P + p) * 6;
    const float inv = 1.f / scale[b];
    float *o = in + i * 4;
    o[0] = q[0] * inv;
    o[1] = q[1] * inv;
    o[2] = q[2] * inv;
    o[3] = q[3] * inv;
    if (tgt) {
        tgt[i * 2] = (q[4] - q[0] - 2.f * q[2]) * inv;
        tgt[i * 2 + 1] = (q[5] - q[1] - 2.f * q[3]) * inv;
    }
}

// Copy (E, S, E_next) of one slice into pair slot `slot`; S is zero outside the box.
__global__ void k_capture(const float2 *E, const float2 *S, const float2 *En, int4 bx, int ng,
                          float *pairs, long long slot) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= ng * ng) return;
    const int x = p % ng, y = p / ng;
    const bool in = x >= bx.x && x <= bx.z && y >= bx.y && y <= bx.w;
    const float2 e = E[p], n = En[p], s = in ? S[p] : make_float2(0, 0);
    float *o = pairs + (slot * ng * ng + p) * 6;
    o[0] = e.x;
    o[1] = e.y;
    o[2] = s.x;
    o[3] = s.y;
    o[4] = n.x;
    o[5] = n.y;
}

// Inference input of beam slices [s0, s0 + B): fields of storage slice
// (ii + first) % ns from fin, sources of beam slice ii.
__global__ void __launch_bounds__(256) k_slice_scale(const float2 *fin, const float2 *src, const int4 *box,
                                                     int ng, int ns, int first, int s0, float *scale) {
    __shared__ double red[256];
    const int b = blockIdx.x, ii = s0 + b;
    const int P = ng * ng;
    const float2 *E = fin + (long long)((ii + first) % ns) * P;
    const float2 *S = src + (long long)ii * P;
    const int4 bx = box[ii];
    double s = 0;
    for (int p = threadIdx.x; p < P; p += 256) {
        const int x = p % ng, y = p / ng;
        const float2 e = E[p];
        s += (double)e.x * e.x + (double)e.y * e.y;
        if (x >= bx.x && x <= bx.z && y >= bx.y && y <= bx.w) {
            const float2 v = S[p];
            s += (double)v.x * v.x + (double)v.y * v.y;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) scale[b] = (float)(sqrt(red[0] / P) + 1e-30);
}

__global__ void k_slice_in(const float2 *fin, const float2 *src, const int4 *box, int ng, int ns, int first,
                           int s0, int B, const float *scale, float *in) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const int P = ng * ng;
    if (i >= (long long)B * P) return;
    const int b = (int)(i / P), p = (int)(i % P), ii = s0 + b;
    const int x = p % ng, y = p / ng;
    const float2 e = fin[(long long)((ii + first) % ns) * P + p];
    const int4 bx = box[ii];
    const float2 sv = (x >= bx.x && x <= bx.z && y >= bx.y && y <= bx.w) ? src[(long long)ii * P + p]
                                                                         : make_float2(0, 0);
    const float inv = 1.f / scale[b];
    float *o = in + i * 4;
    o[0] = e.x * inv;
    o[1] = e.y * inv;
    o[2] = sv.x * inv;
    o[3] = sv.y * inv;
}

// E_next = E + 2 S + scale * out for beam slices [s0, s0 + B) into fout
__global__ void k_slice_out(const float2 *fin, float2 *fout, const float2 *src, const int4 *box, int ng, int ns,
                            int first, int s0, int B, const float *scale, const float *out) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const int P = ng * ng;
    if (i >= (long long)B * P) return;
    const int b = (int)(i / P), p = (int)(i % P), ii = s0 + b;
    const int x = p % ng, y = p / ng;
    const int4 bx = box[ii];
    const float2 sv = (x >= bx.x && x <= bx.z && y >= bx.y && y <= bx.w) ? src[(long long)ii * P + p]
                                                                         : make_float2(0, 0);
    const long long fo = (long long)((ii + first) % ns) * P + p;
    const float2 e = fin[fo];
    fout[fo] = make_float2(e.x + 2.f * sv.x + scale[b] * out[i * 2], e.y + 2.f * sv.y + scale[b] * out[i * 2 + 1]);
}

// Error of the network's field step against the ADI's on one storage slice
// per block: parts[j] = (sum |E_fno - E_adi|^2, sum |E_adi - E|^2,
// sum |E_adi|^2) with E the field before the step.
__global__ void __launch_bounds__(256) k_judge(const float2 *efno, const float2 *eadi, const float2 *ein, int P,
                                               double *parts) {
    __shared__ double red[3][256];
    const long long j = blockIdx.x;
    const float2 *a = efno + j * P, *b = eadi + j * P, *e = ein + j * P;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int p = threadIdx.x; p < P; p += 256) {
        const float2 u = a[p], v = b[p], w = e[p];
        const double dx = (double)u.x - v.x, dy = (double)u.y - v.y;
        const double ux = (double)v.x - w.x, uy = (double)v.y - w.y;
        s0 += dx * dx + dy * dy;
        s1 += ux * ux + uy * uy;
        s2 += (double)v.x * v.x + (double)v.y * v.y;
    }
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    red[2][threadIdx.x] = s2;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x < 3) parts[j * 3 + threadIdx.x] = red[threadIdx.x][0];
}

inline unsigned nb(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void fno_judge_sums(hipStream_t s, const float2 *efno, const float2 *eadi, const float2 *ein, int nslices, int P,
                    double *parts) {
    hipLaunchKernelGGL(k_judge, dim3(nslices), dim3(256), 0, s, efno, eadi, ein, P, parts);
}

namespace {
// waves along one output dimension: the fewest padded rows, ties to the larger tile
int pick_waves(int n, int maxw) {
    int best = 1;
    long long pad = (long long)((n + 31) / 32) * 32;
    for (int w = 2; w <= maxw; w *= 2) {
        const long long p = (long long)((n + 32 * w - 1) / (32 * w)) * 32 * w;
        if (p <= pad) {
            pad = p;
            best = w;
        }
    }
    return best;
}
bool al16(const void *p) { return ((uintptr_t)p & 15) == 0; }
}  // namespace

void launch_gemm_bf16(hipStream_t s, const GemmArgs &g, int batch) {
    const int wm = pick_waves(g.M, 4), wn = pick_waves(g.N, 4 / wm);
    // float4 reads need 16-byte aligned rows and batch bases
    const int vecA = al16(g.A) && g.lda % 4 == 0 && (batch == 1 || g.sA % 4 == 0);
    const int vecB = al16(g.B) && g.ldb % 4 == 0 && (batch == 1 || g.sB % 4 == 0);
    const dim3 grid((unsigned)((g.N + 32 * wn - 1) / (32 * wn)), (unsigned)((g.M + 32 * wm - 1) / (32 * wm)),
                    (unsigned)batch);
    const dim3 blk(64 * wm * wn);
    if (wm == 1 && wn == 1) hipLaunchKernelGGL((k_gemm_bf16<1, 1>), grid, blk, 0, s, g, vecA, vecB);
    else if (wm == 1 && wn == 2) hipLaunchKernelGGL((k_gemm_bf16<1, 2>), grid, blk, 0, s, g, vecA, vecB);
    else if (wm == 1) hipLaunchKernelGGL((k_gemm_bf16<1, 4>), grid, blk, 0, s, g, vecA, vecB);
    else if (wm == 2 && wn == 1) hipLaunchKernelGGL((k_gemm_bf16<2, 1>), grid, blk, 0, s, g, vecA, vecB);
    else if (wm == 2) hipLaunchKernelGGL((k_gemm_bf16<2, 2>), grid, blk, 0, s, g, vecA, vecB);
    else hipLaunchKernelGGL((k_gemm_bf16<4, 1>), grid, blk, 0, s, g, vecA, vecB);
}

// ===========================================================================
// the network
// ===========================================================================
struct Fno {
    int dev = 0;
    hipStream_t s = nullptr;
    FnoShape sh{};
    int N = 0, P = 0, C = 0, L = 0, H = 0, nm = 0, nmode = 0, B = 0;
    // parameter layout (offsets into params)
    long long o_P = 0, o_bP = 0, o_Q1 = 0, o_bQ1 = 0, o_Q2 = 0, o_bQ2 = 0, nparam = 0;
    std::vector<long long> o_Rr, o_Ri, o_W, o_b;
    float *params = nullptr, *grads = nullptr, *am = nullptr, *av = nullptr;
    long long adam_t = 0;
    // constants
    float *Fx = nullptr, *Fy = nullptr, *Gy2 = nullptr, *Gx = nullptr, *Wreal = nullptr, *dWreal = nullptr;
    // activations
    float *in = nullptr, *tgt = nullptr, *scale = nullptr;
    std::vector<float *> a, z, Ah;
    float *X1 = nullptr, *T = nullptr, *Bh = nullptr, *Zb = nullptr, *Y1 = nullptr;
    float *zq = nullptr, *q = nullptr, *out = nullptr;
    float *da = nullptr, *da2 = nullptr, *dq = nullptr, *dout = nullptr, *parts = nullptr, *parts2 = nullptr;
    long long parts_n = 0;
    double *d_loss = nullptr, *colpart = nullptr;
    int *d_idx = nullptr;
    // pair set
    float *pairs = nullptr;
    long long pairs_cap = 0, pairs_n = 0;
    std::vector<void *> allocs;
};

namespace {

int falloc(Fno *f, float **p, long long n) {
    void *q = nullptr;
    if (hipMalloc(&q, (size_t)std::max(n, 1LL) * sizeof(float)) != hipSuccess) {
        set_error("fno: hipMalloc failed");
        return G4MI_E_HIP;
    }
    f->allocs.push_back(q);
    *p = (float *)q;
    return G4MI_OK;
}

// epilogue options of one product
struct Epi {
    const float *bias = nullptr;
    const float *dmul = nullptr;
    float *act_out = nullptr;
    int act = 0;
};

void gemm(Fno *f, const float *A, const float *B, float *C, int M, int N, int K, int lda, int ldb, int ldc,
          long long sA, long long sB, long long sC, int batch, int tA = 0, int tB = 0, float beta = 0.f,
          long long ktot = 0, const Epi &ep = Epi()) {
    GemmArgs g{A, B, C, M, N, K, lda, ldb, ldc, sA, sB, sC, tA, tB, 1.f, beta, ktot,
               ep.bias, ep.dmul, ep.act_out, ep.act};
    launch_gemm_bf16(f->s, g, batch);
}

// weight gradient dW (M x N) = op(A)^T-style product with K = rows (large):
// split K into chunks, partial products into parts, then a sum
constexpr int kWgradK = 256;   // rows per split-K block
constexpr int kWgradG = 32;    // first-stage groups of the partial sum
void gemm_wgrad(Fno *f, const float *A, const float *B, float *dW, int M, int N, long long K, int lda,
                int ldb) {
    const int kc = kWgradK;
    const int nz = (int)((K + kc - 1) / kc);
    const long long n = (long long)M * N;
    gemm(f, A, B, f->parts, M, N, kc, lda, ldb, N, (long long)kc * lda, (long long)kc * ldb, n, nz, 1, 0, 0.f, K);
    const int zc = (nz + kWgradG - 1) / kWgradG, g = (nz + zc - 1) / zc;
    hipLaunchKernelGGL(k_reduce_parts, dim3(nb(n), g), dim3(256), 0, f->s, f->parts, nz, zc, n, f->parts2);
    hipLaunchKernelGGL(k_reduce_parts, dim3(nb(n), 1), dim3(256), 0, f->s, f->parts2, g, g, n, dW);
}

void colsum(Fno *f, const float *d, long long rows, int ch, float *out) {
    hipLaunchKernelGGL(k_colsum_part, dim3(kColBlocks), dim3(256), 0, f->s, d, rows, ch, f->colpart);
    hipLaunchKernelGGL(k_colsum_fin, dim3((ch + 63) / 64), dim3(64), 0, f->s, f->colpart, ch, out);
}

void build_wreal(Fno *f) {
    for (int l = 0; l < f->L; l++)
        hipLaunchKernelGGL(k_build_wreal, dim3(nb((long long)f->nmode * f->C * f->C)), dim3(256), 0, f->s,
                           f->params + f->o_Rr[l], f->params + f->o_Ri[l],
                           f->Wreal + (long long)l * f->nmode * 4 * f->C * f->C, f->nmode, f->C);
}

// forward pass on f->in (B samples) -> f->out
void forward(Fno *f, int B) {
    const int N = f->N, P = f->P, C = f->C, nm = f->nm, nmode = f->nmode, H = f->H;
    const long long rows = (long long)B * P;
    float *prm = f->params;
    // lift
    {
        Epi ep;
        ep.bias = prm + f->o_bP;
        ep.act_out = f->a[0];
        gemm(f, f->in, prm + f->o_P, f->z[0], (int)rows, C, 4, 4, C, C, 0, 0, 0, 1, 0, 0, 0.f, 0, ep);
    }
    for (int l = 0; l < f->L; l++) {
        float *a = f->a[l], *zl = f->z[l + 1];
        // x transform: X1[b][y] (2nm x C) = Fx (2nm x N) a[b][y] (N x C)
        gemm(f, f->Fx, a, f->X1, 2 * nm, C, N, N, C, C, 0, (long long)N * C, 2LL * nm * C, B * N);
        // y transform: T[b] (2nm x 2nm C) = Fy (2nm x N) X1[b] (N x 2nm C)
        gemm(f, f->Fy, f->X1, f->T, 2 * nm, 2 * nm * C, N, N, 2 * nm * C, 2 * nm * C, 0, (long long)N * 2 * nm * C,
             4LL * nm * nm * C, B);
        hipLaunchKernelGGL(k_comb_fwd, dim3(nb((long long)B * nm * nm * C)), dim3(256), 0, f->s, f->T, f->Ah[l], B,
                           nm, C);
        // per-mode channel mixing: Bh[b][mode] = Ah[b][mode] Wreal[mode]
        gemm(f, f->Ah[l], f->Wreal + (long long)l * nmode * 4 * C * C, f->Bh, B, 2 * C, 2 * C, nmode * 2 * C, 2 * C,
             nmode * 2 * C, 2 * C, 4LL * C * C, 2 * C, nmode);
        // inverse y: Z[b] (2N x nm 2C) = Gy2 (2N x nm) Bh[b] (nm x nm 2C)
        gemm(f, f->Gy2, f->Bh, f->Zb, 2 * N, nm * 2 * C, nm, nm, nm * 2 * C, nm * 2 * C, 0, (long long)nmode * 2 * C,
             2LL * N * nm * 2 * C, B);
        hipLaunchKernelGGL(k_comb_inv, dim3(nb((long long)B * N * nm * C)), dim3(256), 0, f->s, f->Zb, f->Y1, B, N,
                           nm, C);
        // inverse x: s[b][y] (N x C) = Gx (N x 2nm) Y1[b][y] (2nm x C)
        gemm(f, f->Gx, f->Y1, zl, N, C, 2 * nm, 2 * nm, C, C, 0, 2LL * nm * C, (long long)N * C, B * N);
        // + a W, + b, activation
        Epi ep;
        ep.bias = prm + f->o_b[l];
        ep.act_out = f->a[l + 1];
        ep.act = l < f->L - 1 ? 1 : 0;
        gemm(f, a, prm + f->o_W[l], zl, (int)rows, C, C, C, C, C, 0, 0, 0, 1, 0, 0, 1.f, 0, ep);
    }
    // projection
    Epi e1;
    e1.bias = prm + f->o_bQ1;
    e1.act_out = f->q;
    e1.act = 1;
    gemm(f, f->a[f->L], prm + f->o_Q1, f->zq, (int)rows, H, C, C, H, H, 0, 0, 0, 1, 0, 0, 0.f, 0, e1);
    Epi e2;
    e2.bias = prm + f->o_bQ2;
    gemm(f, f->q, prm + f->o_Q2, f->out, (int)rows, 2, H, H, 2, 2, 0, 0, 0, 1, 0, 0, 0.f, 0, e2);
}

// backward from f->dout into f->grads
void backward(Fno *f, int B) {
    const int N = f->N, P = f->P, C = f->C, nm = f->nm, nmode = f->nmode, H = f->H;
    const long long rows = (long long)B * P;
    float *prm = f->params, *gr = f->grads;
    // projection
    gemm_wgrad(f, f->q, f->dout, gr + f->o_Q2, H, 2, rows, H, 2);
    colsum(f, f->dout, rows, 2, gr + f->o_bQ2);
    {
        Epi ep;
        ep.dmul = f->zq;
        gemm(f, f->dout, prm + f->o_Q2, f->dq, (int)rows, H, 2, 2, 2, H, 0, 0, 0, 1, 0, 1, 0.f, 0, ep);
    }
    gemm_wgrad(f, f->a[f->L], f->dq, gr + f->o_Q1, C, H, rows, C, H);
    colsum(f, f->dq, rows, H, gr + f->o_bQ1);
    float *dA = f->da, *dA2 = f->da2;
    gemm(f, f->dq, prm + f->o_Q1, dA, (int)rows, C, H, H, H, C, 0, 0, 0, 1, 0, 1);
    for (int l = f->L - 1; l >= 0; l--) {
        float *a = f->a[l];
        // dz = dA * GELU'(z[l+1]), applied by the epilogue of the product
        // that completed dA (no activation after the last layer)
        float *dz = dA;
        colsum(f, dz, rows, C, gr + f->o_b[l]);
        gemm_wgrad(f, a, dz, gr + f->o_W[l], C, C, rows, C, C);
        // dA2 = dz W^T
        gemm(f, dz, prm + f->o_W[l], dA2, (int)rows, C, C, C, C, C, 0, 0, 0, 1, 0, 1);
        // spectral path: dY1[b][y] (2nm x C) = Gx^T (2nm x N) dz[b][y] (N x C)
        gemm(f, f->Gx, dz, f->Y1, 2 * nm, C, N, 2 * nm, C, C, 0, (long long)N * C, 2LL * nm * C, B * N, 1);
        hipLaunchKernelGGL(k_comb_inv_back, dim3(nb((long long)B * N * nm * C)), dim3(256), 0, f->s, f->Y1, f->Zb, B,
                           N, nm, C);
        // dBh[b] (nm x nm 2C) = Gy2^T (nm x 2N) dZ[b] (2N x nm 2C)
        gemm(f, f->Gy2, f->Zb, f->Bh, nm, nm * 2 * C, 2 * N, nm, nm * 2 * C, nm * 2 * C, 0, 2LL * N * nm * 2 * C,
             (long long)nmode * 2 * C, B, 1);
        // dWreal[mode] = Ah[mode]^T dBh[mode]; dAh[mode] = dBh[mode] Wreal[mode]^T
        gemm(f, f->Ah[l], f->Bh, f->dWreal, 2 * C, 2 * C, B, nmode * 2 * C, nmode * 2 * C, 2 * C, 2 * C, 2 * C,
             4LL * C * C, nmode, 1);
        hipLaunchKernelGGL(k_fold_wreal, dim3(nb((long long)nmode * C * C)), dim3(256), 0, f->s, f->dWreal,
                           gr + f->o_Rr[l], gr + f->o_Ri[l], nmode, C);
        float *dAh = f->X1;  // scratch [B][nmode][2C] fits in X1 ([B][N][2nm][C])
        gemm(f, f->Bh, f->Wreal + (long long)l * nmode * 4 * C * C, dAh, B, 2 * C, 2 * C, nmode * 2 * C, 2 * C,
             nmode * 2 * C, 2 * C, 4LL * C * C, 2 * C, nmode, 0, 1);
        hipLaunchKernelGGL(k_comb_fwd_back, dim3(nb((long long)B * nm * nm * C)), dim3(256), 0, f->s, dAh, f->T, B, nm,
                           C);
        // dX1[b] (N x 2nm C) = Fy^T (N x 2nm) dT[b] (2nm x 2nm C)
        gemm(f, f->Fy, f->T, f->X1, N, 2 * nm * C, 2 * nm, N, 2 * nm * C, 2 * nm * C, 0, 4LL * nm * nm * C,
             (long long)N * 2 * nm * C, B, 1);
        // dA2[b][y] (N x C) += Fx^T (N x 2nm) dX1[b][y] (2nm x C)
        Epi ep;
        if (l > 0) ep.dmul = f->z[l];  // a[l] = GELU(z[l]) for 0 < l < L
        gemm(f, f->Fx, f->X1, dA2, N, C, 2 * nm, N, C, C, 0, 2LL * nm * C, (long long)N * C, B * N, 1, 0, 1.f, 0,
             ep);
        std::swap(dA, dA2);
    }
    // lift
    gemm_wgrad(f, f->in, dA, gr + f->o_P, 4, C, rows, 4, C);
    colsum(f, dA, rows, C, gr + f->o_bP);
}

void init_params(Fno *f, unsigned seed) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<float> u01(0.f, 1.f);
    std::vector<float> h((size_t)f->nparam, 0.f);
    auto fill = [&](long long off, long long n, float lim) {
        for (long long i = 0; i < n; i++) h[(size_t)(off + i)] = (2.f * u01(rng) - 1.f) * lim;
    };
    const int C = f->C, H = f->H;
    fill(f->o_P, 4LL * C, 1.f / std::sqrt(4.f));
    for (int l = 0; l < f->L; l++) {
        const float sc = 1.f / (float)(C * C);
        for (long long i = 0; i < (long long)f->nmode * C * C; i++) {
            h[(size_t)(f->o_Rr[l] + i)] = sc * u01(rng);
            h[(size_t)(f->o_Ri[l] + i)] = sc * u01(rng);
        }
        fill(f->o_W[l], (long long)C * C, 1.f / std::sqrt((float)C));
    }
    fill(f->o_Q1, (long long)C * H, 1.f / std::sqrt((float)C));
    fill(f->o_Q2, (long long)H * 2, 1.f / std::sqrt((float)H));
    (void)hipMemcpy(f->params, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice);
}

}  // namespace

// ---------------------------------------------------------------------------
// internal API (used by the C-ABI below and by the context's deployment)
// ---------------------------------------------------------------------------
Fno *fno_new(int device, const FnoShape &sh, unsigned seed) {
    if (sh.ngrid < 8 || sh.width < 1 || sh.width > 256 || sh.modes < 2 || sh.modes % 2 || sh.modes > sh.ngrid ||
        sh.layers < 1 || sh.hidden < 1 || sh.hidden > 256 || sh.max_batch < 1) {
        set_error("fno: invalid shape");
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) {
        set_error("fno: no usable HIP device");
        return nullptr;
    }
    Fno *f = new Fno;
    f->dev = device;
    f->sh = sh;
    f->N = sh.ngrid;
    f->P = sh.ngrid * sh.ngrid;
    f->C = sh.width;
    f->L = sh.layers;
    f->H = sh.hidden;
    f->nm = sh.modes;
    f->nmode = sh.modes * sh.modes;
    f->B = sh.max_batch;
    const int C = f->C, N = f->N, nm = f->nm, B = f->B, H = f->H;
    const long long P = f->P;
    long long o = 0;
    f->o_P = o; o += 4LL * C;
    f->o_bP = o; o += C;
    for (int l = 0; l < f->L; l++) {
        f->o_Rr.push_back(o); o += (long long)f->nmode * C * C;
        f->o_Ri.push_back(o); o += (long long)f->nmode * C * C;
        f->o_W.push_back(o); o += (long long)C * C;
        f->o_b.push_back(o); o += C;
    }
    f->o_Q1 = o; o += (long long)C * H;
    f->o_bQ1 = o; o += H;
    f->o_Q2 = o; o += 2LL * H;
    f->o_bQ2 = o; o += 2;
    f->nparam = o;
    auto fail = [&]() {
        fno_free(f);
        return (Fno *)nullptr;
    };
    if (hipStreamCreateWithFlags(&f->s, hipStreamNonBlocking) != hipSuccess) return fail();
    const long long act = (long long)B * P * C;
    f->a.resize(f->L + 1);
    f->z.resize(f->L + 1);
    f->Ah.resize(f->L);
    int rc = 0;
    rc |= falloc(f, &f->params, o) | falloc(f, &f->grads, o) | falloc(f, &f->am, o) | falloc(f, &f->av, o);
    rc |= falloc(f, &f->Fx, 2LL * nm * N) | falloc(f, &f->Fy, 2LL * nm * N) | falloc(f, &f->Gy2, 2LL * N * nm) |
          falloc(f, &f->Gx, (long long)N * 2 * nm);
    rc |= falloc(f, &f->Wreal, (long long)f->L * f->nmode * 4 * C * C) |
          falloc(f, &f->dWreal, (long long)f->nmode * 4 * C * C);
    rc |= falloc(f, &f->in, (long long)B * P * 4) | falloc(f, &f->tgt, (long long)B * P * 2) |
          falloc(f, &f->scale, B);
    for (int l = 0; l <= f->L; l++) rc |= falloc(f, &f->a[l], act) | falloc(f, &f->z[l], act);
    for (int l = 0; l < f->L; l++) rc |= falloc(f, &f->Ah[l], (long long)B * f->nmode * 2 * C);
    rc |= falloc(f, &f->X1, (long long)B * N * 2 * nm * C) | falloc(f, &f->T, (long long)B * 4 * nm * nm * C) |
          falloc(f, &f->Bh, (long long)B * f->nmode * 2 * C) | falloc(f, &f->Zb, (long long)B * 2 * N * nm * 2 * C) |
          falloc(f, &f->Y1, (long long)B * N * 2 * nm * C);
    rc |= falloc(f, &f->zq, (long long)B * P * H) | falloc(f, &f->q, (long long)B * P * H) |
          falloc(f, &f->out, (long long)B * P * 2);
    rc |= falloc(f, &f->da, act) | falloc(f, &f->da2, act) | falloc(f, &f->dq, (long long)B * P * H) |
          falloc(f, &f->dout, (long long)B * P * 2);
    const long long rows = (long long)B * P;
    const long long wmax = std::max({C * C, C * H, 2 * H, 4 * C});
    f->parts_n = ((rows + kWgradK - 1) / kWgradK) * wmax;
    rc |= falloc(f, &f->parts, f->parts_n);
    rc |= falloc(f, &f->parts2, kWgradG * wmax);
    if (rc) return fail();
    if (hipMalloc(&f->d_loss, sizeof(double)) != hipSuccess) return fail();
    f->allocs.push_back(f->d_loss);
    if (hipMalloc(&f->d_idx, (size_t)B * sizeof(int)) != hipSuccess) return fail();
    f->allocs.push_back(f->d_idx);
    if (hipMalloc(&f->colpart, (size_t)kColBlocks * std::max(C, H) * sizeof(double)) != hipSuccess) return fail();
    f->allocs.push_back(f->colpart);
    // DFT matrices over the retained modes k = -nm/2 .. nm/2 - 1
    std::vector<float> fx(2 * nm * N), gy2(2 * N * nm), gx(N * 2 * nm);
    const double tw = 2.0 * M_PI / N, inv = 1.0 / ((double)N * N);
    for (int r = 0; r < nm; r++) {
        const int k = r - nm / 2;
        for (int x = 0; x < N; x++) {
            const double th = tw * k * x;
            fx[(size_t)r * N + x] = (float)cos(th);
            fx[(size_t)(nm + r) * N + x] = (float)(-sin(th));
            gy2[(size_t)x * nm + r] = (float)cos(th);
            gy2[(size_t)(N + x) * nm + r] = (float)sin(th);
            gx[(size_t)x * 2 * nm + r] = (float)(cos(th) * inv);
            gx[(size_t)x * 2 * nm + nm + r] = (float)(-sin(th) * inv);
        }
    }
    (void)hipMemcpy(f->Fx, fx.data(), fx.size() * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(f->Fy, fx.data(), fx.size() * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(f->Gy2, gy2.data(), gy2.size() * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(f->Gx, gx.data(), gx.size() * 4, hipMemcpyHostToDevice);
    init_params(f, seed);
    (void)hipMemset(f->am, 0, o * sizeof(float));
    (void)hipMemset(f->av, 0, o * sizeof(float));
    build_wreal(f);
    if (hipStreamSynchronize(f->s) != hipSuccess) return fail();
    return f;
}

void fno_free(Fno *f) {
    if (!f) return;
    if (f->s) (void)hipStreamSynchronize(f->s);
    for (void *p : f->allocs) (void)hipFree(p);
    if (f->pairs) (void)hipFree(f->pairs);
    if (f->s) (void)hipStreamDestroy(f->s);
    delete f;
}

long long fno_nparam(const Fno *f) { return f->nparam; }
int fno_ngrid(const Fno *f) { return f->N; }
int fno_max_batch(const Fno *f) { return f->B; }
long long fno_pairs(const Fno *f) { return f->pairs_n; }

int fno_get_params(Fno *f, float *h) {
    G4_HIP(hipMemcpyAsync(h, f->params, f->nparam * sizeof(float), hipMemcpyDeviceToHost, f->s));
    G4_HIP(hipStreamSynchronize(f->s));
    return G4MI_OK;
}

int fno_set_params(Fno *f, const float *h) {
    G4_HIP(hipMemcpyAsync(f->params, h, f->nparam * sizeof(float), hipMemcpyHostToDevice, f->s));
    build_wreal(f);
    G4_HIP(hipStreamSynchronize(f->s));
    return G4MI_OK;
}

int fno_forward_host(Fno *f, const float *in, float *out, int B) {
    if (B < 1 || B > f->B) {
        set_error("fno: batch out of range");
        return G4MI_E_ARG;
    }
    G4_HIP(hipMemcpyAsync(f->in, in, (size_t)B * f->P * 4 * sizeof(float), hipMemcpyHostToDevice, f->s));
    forward(f, B);
    G4_HIP(hipMemcpyAsync(out, f->out, (size_t)B * f->P * 2 * sizeof(float), hipMemcpyDeviceToHost, f->s));
    G4_HIP(hipStreamSynchronize(f->s));
    G4_HIP(hipGetLastError());
    return G4MI_OK;
}

int fno_loss_grad_host(Fno *f, const float *in, const float *tgt, int B, double *loss, float *grad) {
    if (B < 1 || B > f->B) {
        set_error("fno: batch out of range");
        return G4MI_E_ARG;
    }
    G4_HIP(hipMemcpyAsync(f->in, in, (size_t)B * f->P * 4 * sizeof(float), hipMemcpyHostToDevice, f->s));
    G4_HIP(hipMemcpyAsync(f->tgt, tgt, (size_t)B * f->P * 2 * sizeof(float), hipMemcpyHostToDevice, f->s));
    forward(f, B);
    G4_HIP(hipMemsetAsync(f->d_loss, 0, sizeof(double), f->s));
    const long long n = (long long)B * f->P * 2;
    hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, f->s, f->out, f->tgt, f->dout, n, f->d_loss);
    backward(f, B);
    G4_HIP(hipMemcpyAsync(loss, f->d_loss, sizeof(double), hipMemcpyDeviceToHost, f->s));
    if (grad) G4_HIP(hipMemcpyAsync(grad, f->grads, f->nparam * sizeof(float), hipMemcpyDeviceToHost, f->s));
    G4_HIP(hipStreamSynchronize(f->s));
    G4_HIP(hipGetLastError());
    return G4MI_OK;
}

int fno_reserve(Fno *f, long long npairs) {
    if (npairs <= f->pairs_cap) return G4MI_OK;
    float *p = nullptr;
    G4_HIP(hipMalloc(&p, (size_t)npairs * f->P * 6 * sizeof(float)));
    if (f->pairs_n > 0)
        G4_HIP(hipMemcpy(p, f->pairs, (size_t)f->pairs_n * f->P * 6 * sizeof(float), hipMemcpyDeviceToDevice));
    if (f->pairs) (void)hipFree(f->pairs);
    f->pairs = p;
    f->pairs_cap = npairs;
    return G4MI_OK;
}

int fno_capture(Fno *f, hipStream_t s, const float2 *E, const float2 *S, const float2 *En, int4 box) {
    if (f->pairs_n >= f->pairs_cap) {
        set_error("fno: pair set full");
        return G4MI_E_STATE;
    }
    hipLaunchKernelGGL(k_capture, dim3(nb(f->P)), dim3(256), 0, s, E, S, En, box, f->N, f->pairs, f->pairs_n);
    f->pairs_n++;
    return G4MI_OK;
}

// One pass over the pairs [p0, p1) in batches (shuffled when train), Adam
// steps when train; returns the mean loss.
static int run_pairs(Fno *f, long long p0, long long p1, int batch, bool train, float lr, std::mt19937_64 &rng,
                     double *mean_loss) {
    std::vector<int> order;
    for (long long i = p0; i < p1; i++) order.push_back((int)i);
    if (train) std::shuffle(order.begin(), order.end(), rng);
    double tot = 0;
    long long nb_ = 0;
    for (size_t k = 0; k < order.size(); k += (size_t)batch) {
        const int B = (int)std::min((size_t)batch, order.size() - k);
        G4_HIP(hipMemcpyAsync(f->d_idx, order.data() + k, B * sizeof(int), hipMemcpyHostToDevice, f->s));
        hipLaunchKernelGGL(k_scale, dim3(B), dim3(256), 0, f->s, f->pairs, f->d_idx, f->P, f->scale);
        hipLaunchKernelGGL(k_assemble, dim3(nb((long long)B * f->P)), dim3(256), 0, f->s, f->pairs, f->d_idx,
                           f->scale, B, f->P, f->in, f->tgt);
        forward(f, B);
        G4_HIP(hipMemsetAsync(f->d_loss, 0, sizeof(double), f->s));
        const long long n = (long long)B * f->P * 2;
        hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, f->s, f->out, f->tgt, f->dout, n, f->d_loss);
        if (train) {
            backward(f, B);
            f->adam_t++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - std::pow(b1, (float)f->adam_t), c2 = 1.f - std::pow(b2, (float)f->adam_t);
            hipLaunchKernelGGL(k_adam, dim3(nb(f->nparam)), dim3(256), 0, f->s, f->params, f->grads, f->am, f->av,
                               f->nparam, lr, b1, b2, 1e-8f, c1, c2);
            build_wreal(f);
        }
        double l = 0;
        G4_HIP(hipMemcpyAsync(&l, f->d_loss, sizeof(double), hipMemcpyDeviceToHost, f->s));
        G4_HIP(hipStreamSynchronize(f->s));
        tot += l * B;
        nb_ += B;
    }
    G4_HIP(hipGetLastError());
    *mean_loss = nb_ ? tot / (double)nb_ : 0.0;
    return G4MI_OK;
}

int fno_train_span(Fno *f, int first_epoch, int epochs, int total_epochs, int batch, float lr,
                   long long nholdout, unsigned seed, double *train_loss, double *holdout_loss) {
    if (batch < 1 || batch > f->B || epochs < 0 || first_epoch < 0 || total_epochs < first_epoch + epochs ||
        nholdout < 0 || nholdout >= f->pairs_n) {
        set_error("fno_train: invalid arguments (batch, epochs or hold-out size)");
        return G4MI_E_ARG;
    }
    const long long ntrain = f->pairs_n - nholdout;
    for (int k = 0; k < epochs; k++) {
        const int e = first_epoch + k;
        std::mt19937_64 rng(seed + 0x9E3779B97F4A7C15ULL * (unsigned long long)(e + 1));
        // cosine schedule over all total_epochs
        const float lre = lr * 0.5f * (1.f + std::cos((float)M_PI * (float)e / (float)std::max(1, total_epochs)));
        double l = 0;
        G4_TRY(run_pairs(f, 0, ntrain, batch, true, lre, rng, &l));
        if (train_loss) train_loss[k] = l;
        if (holdout_loss) {
            double h = 0;
            if (nholdout > 0) G4_TRY(run_pairs(f, ntrain, f->pairs_n, batch, false, 0.f, rng, &h));
            holdout_loss[k] = h;
        }
    }
    return G4MI_OK;
}

int fno_train(Fno *f, int epochs, int batch, float lr, long long nholdout, unsigned seed, double *train_loss,
              double *holdout_loss) {
    return fno_train_span(f, 0, epochs, epochs, batch, lr, nholdout, seed, train_loss, holdout_loss);
}

int fno_eval(Fno *f, long long p0, long long p1, int batch, double *loss) {
    std::mt19937_64 rng(1);
    if (p0 < 0 || p1 > f->pairs_n || p0 >= p1 || batch < 1 || batch > f->B) {
        set_error("fno_eval: invalid range");
        return G4MI_E_ARG;
    }
    return run_pairs(f, p0, p1, batch, false, 0.f, rng, loss);
}

// Batched inference over beam slices: fout[(ii + first) % ns] = fin[...] +
// 2 src[ii] + scale * FNO((fin, src) / scale) for every slice ii.
int fno_infer_slices(Fno *f, hipStream_t s, const float2 *fin, float2 *fout, const float2 *src, const int4 *box,
                     int ns, int first) {
    hipStream_t keep = f->s;
    f->s = s;
    for (int s0 = 0; s0 < ns; s0 += f->B) {
        const int B = std::min(f->B, ns - s0);
        hipLaunchKernelGGL(k_slice_scale, dim3(B), dim3(256), 0, s, fin, src, box, f->N, ns, first, s0, f->scale);
        hipLaunchKernelGGL(k_slice_in, dim3(nb((long long)B * f->P)), dim3(256), 0, s, fin, src, box, f->N, ns, first,
                           s0, B, f->scale, f->in);
        forward(f, B);
        hipLaunchKernelGGL(k_slice_out, dim3(nb((long long)B * f->P)), dim3(256), 0, s, fin, fout, src, box, f->N,
                           ns, first, s0, B, f->scale, f->out);
    }
    f->s = keep;
    G4_HIP(hipGetLastError());
    return G4MI_OK;
}

}  // namespace g4mi

using namespace g4mi;

struct g4mi_fno {
    Fno *f;
};

extern "C" {

int g4mi_fno_create(g4mi_fno **out, int device, int ngrid, int width, int modes, int layers, int hidden,
                    int max_batch, unsigned seed) {
    if (!out) return G4MI_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1 || device < 0 || device >= ndev) {
        set_error("g4mi_fno_create: no usable HIP device");
        return G4MI_E_NODEV;
    }
    Fno *f = fno_new(device, FnoShape{ngrid, width, modes, layers, hidden, max_batch}, seed);
    if (!f) return G4MI_E_ARG;
    *out = new g4mi_fno{f};
    return G4MI_OK;
}

int g4mi_fno_destroy(g4mi_fno *h) {
    if (!h) return G4MI_OK;
    fno_free(h->f);
    delete h;
    return G4MI_OK;
}

long long g4mi_fno_param_count(g4mi_fno *h) { return h ? fno_nparam(h->f) : G4MI_E_ARG; }
int g4mi_fno_get_params(g4mi_fno *h, float *p) { return (h && p) ? fno_get_params(h->f, p) : G4MI_E_ARG; }
int g4mi_fno_set_params(g4mi_fno *h, const float *p) { return (h && p) ? fno_set_params(h->f, p) : G4MI_E_ARG; }
int g4mi_fno_forward(g4mi_fno *h, const float *in, float *out, int batch) {
    return (h && in && out) ? fno_forward_host(h->f, in, out, batch) : G4MI_E_ARG;
}
int g4mi_fno_loss_grad(g4mi_fno *h, const float *in, const float *tgt, int batch, double *loss, float *grad) {
    return (h && in && tgt && loss) ? fno_loss_grad_host(h->f, in, tgt, batch, loss, grad) : G4MI_E_ARG;
}
int g4mi_fno_reserve(g4mi_fno *h, long long n) { return (h && n >= 0) ? fno_reserve(h->f, n) : G4MI_E_ARG; }
long long g4mi_fno_pair_count(g4mi_fno *h) { return h ? fno_pairs(h->f) : G4MI_E_ARG; }
int g4mi_fno_train(g4mi_fno *h, int epochs, int batch, float lr, long long nholdout, unsigned seed, double *tl,
                   double *hl) {
    return h ? fno_train(h->f, epochs, batch, lr, nholdout, seed, tl, hl) : G4MI_E_ARG;
}
int g4mi_fno_train_span(g4mi_fno *h, int first_epoch, int epochs, int total_epochs, int batch, float lr,
                        long long nholdout, unsigned seed, double *tl, double *hl) {
    return h ? fno_train_span(h->f, first_epoch, epochs, total_epochs, batch, lr, nholdout, seed, tl, hl)
             : G4MI_E_ARG;
}
int g4mi_fno_eval(g4mi_fno *h, long long first, long long last, int batch, double *loss) {
    return (h && loss) ? fno_eval(h->f, first, last, batch, loss) : G4MI_E_ARG;
}

}  // extern "C"

namespace g4mi {
Fno *fno_of(g4mi_fno *h) { return h ? h->f : nullptr; }
}  // namespace g4mi
