// Element-wise losses shared by the loss kernels and the fused FP16 training kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "train.h"

namespace nlod {

// Per-element value and scaled gradient, as the reference's loss kernels
// (losses/binary_cross_entropy.h:7-43, losses/relative_l2.h:39-76,
// losses/l2.h) for one prediction p and target t; n_total = N * dims.
__device__ __forceinline__ double loss_value(int type, float p, float t, float n_total) {
    if (type == LOSS_BCE) {
        const float eps = 1e-5f;
        const float pc = fminf(fmaxf(p, eps), 1.0f - eps);
        // as the reference evaluates it: t*log(p) in float, the (1.0 - t) term in double
        return -((double)(t * logf(pc)) + (1.0 - (double)t) * log(1.0 - (double)pc)) / (double)n_total;
    }
    if (type == LOSS_RELATIVE_L2) {
        const float den = p * p + 0.01f;
        const float d = p - t;
        return (double)(d * d / den / n_total);
    }
    const float d = p - t;  // L2
    return (double)(d * d / n_total);
}

__device__ __forceinline__ float loss_grad(int type, float p, float t, float n_total, float scale) {
    if (type == LOSS_BCE) {
        const float eps = 1e-5f;
        const float pc = fminf(fmaxf(p, eps), 1.0f - eps);
        const float g = -(t - pc) / (pc - pc * pc);
        return scale * g / n_total;
    }
    if (type == LOSS_RELATIVE_L2) {
        const float den = p * p + 0.01f;
        const float d = p - t;
        return scale * (2.0f * d / den) / n_total;
    }
    const float d = p - t;  // L2
    return scale * (2.0f * d) / n_total;
}

__device__ __forceinline__ void loss_elem(int type, float p, float t, float n_total, float scale, double& value,
                                          float& grad) {
    value = loss_value(type, p, t, n_total);
    grad = loss_grad(type, p, t, n_total, scale);
}

}  // namespace nlod
