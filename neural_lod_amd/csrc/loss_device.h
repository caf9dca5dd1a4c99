// Element-wise losses shared by the loss kernels and the fused FP16 training kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "train.h"

namespace nlod {

// Per-element value and scaled gradient, as the reference's loss kernels
// (losses/binary_cross_entropy.h:7-43, losses/relative_l2.h:39-76,
// losses/l2.h) for one prediction p and target t; n_total = N * dims.
__device__ __forceinline__ void loss_elem(int type, float p, float t, float n_total, float scale, double& value,
                                          float& grad) {
    if (type == LOSS_BCE) {
        const float eps = 1e-5f;
        const float pc = fminf(fmaxf(p, eps), 1.0f - eps);
        // as the reference evaluates it: t*log(p) in float, the (1.0 - t) term in double
        value = -((double)(t * logf(pc)) + (1.0 - (double)t) * log(1.0 - (double)pc)) / (double)n_total;
        const float g = -(t - pc) / (pc - pc * pc);
        grad = scale * g / n_total;
    } else if (type == LOSS_RELATIVE_L2) {
        const float den = p * p + 0.01f;
        const float d = p - t;
        value = (double)(d * d / den / n_total);
        grad = scale * (2.0f * d / den) / n_total;
    } else {  // L2
        const float d = p - t;
        value = (double)(d * d / n_total);
        grad = scale * (2.0f * d) / n_total;
    }
}

}  // namespace nlod
