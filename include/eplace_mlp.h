/*
 * eplace_mlp.h -- step-length model: a 16 -> 64 -> 64 -> 1 GELU MLP that
 * proposes the Nesterov step length (log10) from per-iteration features.
 * Training, gradients and inference run in hand-written HIP kernels.
 * No counterpart in the reference (its step is fixed: AIEplace.cpp:1020-1040).
 *
 * Weight blob layout (EP_MLP_NWEIGHTS floats):
 *   [0,16)   feature mean      [16,32)  feature std
 *   32       target mean       33       target std
 *   W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3
 * The model sees x' = (x - mean)/std and predicts t' ; log10(step / bin width) =
 * t' * std_t + mean_t, so the proposed step is bin width * 10^(t' std_t + mean_t).
 * Loss: mean squared error of t' against the normalised target.
 *
 * Features (per iteration, in this order; lengths in bin widths bw, norms per
 * cell: |.| / sqrt(n), n = movable + filler cells, so that instances of any
 * size and grid share one scale): log10(alpha_Lipschitz / bw), log10 of the
 * two previous accepted steps / bw; overflow; overflow change;
 * log10(lambda / lambda_0); log10(gamma / bw); relative HPWL change;
 * log10(|g| / sqrt n); log10(|dv| / (sqrt n bw)); log10(|dg| / sqrt n);
 * Nesterov coefficient; evaluations of the last iteration;
 * log10(lambda |grad D|_1 / |grad W|_1); log10(last alpha_hat / alpha);
 * iteration/1000.
 */
#ifndef EPLACE_MLP_H
#define EPLACE_MLP_H

#include <stdint.h>

#include "eplace.h"

#ifdef __cplusplus
extern "C" {
#endif

#define EP_MLP_IN 16
#define EP_MLP_HID 64
#define EP_MLP_NWEIGHTS (34 + EP_MLP_HID * EP_MLP_IN + EP_MLP_HID + EP_MLP_HID * EP_MLP_HID + EP_MLP_HID + EP_MLP_HID + 1)

typedef struct ep_mlp_train_params {
    int32_t epochs;
    int32_t batch;
    float lr;            /* Adam */
    float beta1, beta2, eps;
    uint64_t seed;       /* weight init and shuffling */
} ep_mlp_train_params;

void ep_mlp_default_train_params(ep_mlp_train_params *p);
/* Trains from scratch on n records; weights receives the blob (normalisation
 * from the records).  loss_curve (nullable, `epochs` entries) gets the mean
 * training loss of each epoch. */
ep_status ep_mlp_train(const float *features, const float *log_step, int32_t n,
                       const ep_mlp_train_params *p, float *weights, float *loss_curve);
/* Inference for n feature rows: out = predicted log10(step / bin width). */
ep_status ep_mlp_forward(const float *weights, const float *features, int32_t n, float *out);
/* Loss and its gradient w.r.t. the weight blob (normalisation entries get 0). */
ep_status ep_mlp_grad(const float *weights, const float *features, const float *log_step, int32_t n,
                      float *grad, float *loss);

#ifdef __cplusplus
}
#endif
#endif
