// ep_mlp.h -- the step-length MLP's weight layout and its single-sample
// forward pass, shared by the proposal kernel (k_mlp.hip) and the step-decision
// block of k_finalize (k_place.hip), so both give bit-identical steps.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/eplace_mlp.h"
#include "ep_ctx.h"

namespace ep {

constexpr int kMlpIn = EP_MLP_IN, kMlpHid = EP_MLP_HID;
// blob: [0,16) feature means, [16,32) feature stds, 32/33 target mean/std, then
// W1 [hid][in], b1, W2 [hid][hid], b2, W3 [hid], b3
constexpr int kMlpW1 = 34, kMlpB1 = kMlpW1 + kMlpHid * kMlpIn, kMlpW2 = kMlpB1 + kMlpHid;
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid, kMlpW3 = kMlpB2 + kMlpHid, kMlpB3 = kMlpW3 + kMlpHid;
static_assert(kMlpB3 + 1 == EP_MLP_NWEIGHTS, "weight blob layout");

__device__ inline float mlp_gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

// The model's step for the features `feat` (16), written to sc->alpha when
// finite and positive (else the Lipschitz step stays).  Every thread of the
// block calls it (it has block barriers); threads 0..kMlpHid-1 compute, and
// only when `run` (uniform over the block) is set.  Thread 0 writes sc->alpha.
__device__ inline void mlp_propose_block(const float *__restrict__ w, const float *feat,
                                         DevScalars *__restrict__ sc, bool run)
{
    __shared__ float xn[kMlpIn], a1[kMlpHid], a2[kMlpHid];
    const int j = threadIdx.x;
    if (run && j < kMlpIn) xn[j] = (feat[j] - w[j]) / w[16 + j];
    __syncthreads();
    if (run && j < kMlpHid) {
        float acc = w[kMlpB1 + j];
        for (int i = 0; i < kMlpIn; ++i) acc += w[kMlpW1 + j * kMlpIn + i] * xn[i];
        a1[j] = mlp_gelu(acc);
    }
    __syncthreads();
    if (run && j < kMlpHid) {
        float acc = w[kMlpB2 + j];
        for (int k = 0; k < kMlpHid; ++k) acc += w[kMlpW2 + j * kMlpHid + k] * a1[k];
        a2[j] = mlp_gelu(acc);
    }
    __syncthreads();
    if (run && j == 0) {
        float o = w[kMlpB3];
        for (int k = 0; k < kMlpHid; ++k) o += w[kMlpW3 + k] * a2[k];
        const float alpha = exp10f(o * w[33] + w[32]) * sc->step_unit;
        if (isfinite(alpha) && alpha > 0.0f) sc->alpha = alpha;
    }
}

}  // namespace ep
