// Fourier-neural-operator surrogate of the per-slice field step.
//
// The network maps (E, S) -> E_next on one slice's ngrid x ngrid grid: the
// field and deposited source of an integration step to the field after
// it, the map the ADI solver computes.  Architecture (FNO2d): a 1x1 lift
// of the 4 input channels (Re/Im of E and S, divided by a per-sample scale)
// to `width` channels, `layers` Fourier layers
//     z = IDFT(R . DFT(a)) + a W + b,   a <- GELU(z) (no GELU after the last),
// each keeping modes x modes Fourier modes (k = -modes/2 .. modes/2-1 per
// axis) with complex width x width weights per mode, and a projection
// width -> hidden -> 2 with GELU between.  The output is the update
// (E_next - E) / scale.
//
// Every dense product (the truncated DFTs, the per-mode channel mixing and
// the 1x1 convolutions, forward and backward) runs on one hand-written
// BF16 MFMA GEMM kernel with FP32 accumulation; master weights, Adam state
// and activations are FP32.
#pragma once

#include "g4mi_common.h"

namespace g4mi {

struct GemmArgs {
    const float *A, *B;
    float *C;
    int M, N, K, lda, ldb, ldc;
    long long sA, sB, sC;   // batch strides (elements)
    int transA, transB;     // A stored K x M / B stored N x K
    float alpha, beta;
    long long ktot;         // > 0: split-K, batch z covers k in [z K, min((z+1) K, ktot))
    // epilogue (all optional): v = alpha acc + beta C (+ bias[col]);
    // v *= GELU'(dmul) (dmul indexed like C); C = v; act_out = GELU(v) when
    // act, else v (act_out indexed like C)
    const float *bias = nullptr;
    const float *dmul = nullptr;
    float *act_out = nullptr;
    int act = 0;
};

// C[z] = alpha op(A[z]) op(B[z]) + beta C[z] for z < batch, with the
// epilogue above.  The output tile (32 WM x 32 WN) is chosen per call from
// M and N.
void launch_gemm_bf16(hipStream_t s, const GemmArgs &g, int batch);

struct Fno;

struct FnoShape {
    int ngrid, width, modes, layers, hidden, max_batch;
};

}  // namespace g4mi
