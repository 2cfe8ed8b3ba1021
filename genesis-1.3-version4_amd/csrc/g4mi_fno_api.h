// Internal API of the FNO surrogate (g4mi_fno.hip) for the C-ABI and for
// a context's capture and deployment paths.
#pragma once

#include "g4mi_fno.h"

namespace g4mi {

Fno *fno_new(int device, const FnoShape &sh, unsigned seed);
void fno_free(Fno *f);
long long fno_nparam(const Fno *f);
int fno_ngrid(const Fno *f);
int fno_max_batch(const Fno *f);
long long fno_pairs(const Fno *f);
int fno_get_params(Fno *f, float *h);
int fno_set_params(Fno *f, const float *h);
// in [B][ngrid^2][4] (E, S already divided by the sample scale), out [B][ngrid^2][2]
int fno_forward_host(Fno *f, const float *in, float *out, int B);
// mean-squared loss of the output against target [B][ngrid^2][2] and the
// gradient with respect to every parameter (parameter order of get_params)
int fno_loss_grad_host(Fno *f, const float *in, const float *tgt, int B, double *loss, float *grad);
int fno_reserve(Fno *f, long long npairs);
// append one (E, S, E_next) pair (device pointers to one slice; S zero outside box)
int fno_capture(Fno *f, hipStream_t s, const float2 *E, const float2 *S, const float2 *En, int4 box);
int fno_train(Fno *f, int epochs, int batch, float lr, long long nholdout, unsigned seed, double *train_loss,
              double *holdout_loss);
int fno_train_span(Fno *f, int first_epoch, int epochs, int total_epochs, int batch, float lr,
                   long long nholdout, unsigned seed, double *train_loss, double *holdout_loss);
int fno_eval(Fno *f, long long p0, long long p1, int batch, double *loss);
// per storage slice j < nslices: parts[3 j ..] = (sum |efno - eadi|^2,
// sum |eadi - ein|^2, sum |eadi|^2) over its P cells
void fno_judge_sums(hipStream_t s, const float2 *efno, const float2 *eadi, const float2 *ein, int nslices, int P,
                    double *parts);
int fno_infer_slices(Fno *f, hipStream_t s, const float2 *fin, float2 *fout, const float2 *src, const int4 *box,
                     int ns, int first);
Fno *fno_of(g4mi_fno *h);

}  // namespace g4mi
