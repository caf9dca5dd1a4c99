// Training step: device-side state and the launchers of train.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "nlod_common.h"

namespace nlod {

enum LossType : int { LOSS_L2 = 0, LOSS_RELATIVE_L2 = 1, LOSS_BCE = 2 };

// Partial-sum slots of the loss and gradient reductions (fixed: deterministic order).
constexpr uint32_t kLossPartials = 512;
constexpr uint32_t kGradPartials = 2048;

// Lives in device memory; read and written only by the step kernels.
struct TrainState {
    float loss_scale;     // scale the next loss evaluation multiplies dL/dy by
    float applied_scale;  // scale of the gradients of the current step
    float lr;             // learning rate of the current step (decay schedule applied)
    float lr_factor;      // ExponentialDecay factor
    uint32_t step;        // completed optimizer steps
    uint32_t good_steps;  // finite steps since the last loss-scale change
    uint32_t found_inf;   // current step's gradients were not finite: step skipped
    uint32_t skipped;     // skipped steps in total
    double loss;          // loss of the current step (sum of per-element values, FP64)
    double grad_sq;       // squared L2 norm of the unscaled gradient (FP64)
};

// Scalars of the decisions taken in finalize_kernel.
struct TrainHyper {
    float base_lr;
    float decay_base;  // ExponentialDecay (decay_interval 0: no schedule)
    uint32_t decay_start, decay_interval, decay_end;
    int dynamic_scale;
    float growth, backoff, min_scale, max_scale;
    uint32_t growth_interval;
};

// Adam hyperparameters (adam.h:214-262); lr comes from TrainState.
struct AdamParams {
    float beta1, beta2, epsilon, l2_reg;
    float relative_decay, absolute_decay, clipping, non_matrix_lr_factor;
    int optimize_matrix, optimize_non_matrix, adabound;
};

void train_state_init(hipStream_t s, TrainState* st, float loss_scale);
// d_out = scale * dL/dpred over n_elems = N * dims elements; FP64 partial loss sums [kLossPartials]
void loss_forward(hipStream_t s, int type, uint64_t n_elems, uint32_t dims, const TrainState* st, const float* pred,
                  const float* target, float* d_out, double* partial);
// per-element values (nullable) and gradients with an explicit scale
void loss_evaluate(hipStream_t s, int type, uint64_t n_elems, float scale, const float* pred, const float* target,
                   float* values, float* grads);
// FP64 partial sums of g^2 and non-finite counts [kGradPartials]
void grad_stats(hipStream_t s, uint64_t n, const float* g, double* partial_sq, uint32_t* partial_bad);
void train_finalize(hipStream_t s, const TrainHyper& h, uint32_t n_loss_partials, const double* loss_partial,
                    const double* grad_partial, const uint32_t* bad_partial, TrainState* st);
// Adam over n parameters (the first n_matrix are matrix weights); w_half (nullable)
// receives an FP16 copy of every updated weight; clear_grads zeroes consumed gradients.
void adam_update(hipStream_t s, const AdamParams& a, uint64_t n, uint64_t n_matrix, const TrainState* st, float* w,
                 void* w_half, float* g, float* m1, float* m2, uint32_t* steps, bool clear_grads);
void ema_update(hipStream_t s, float decay, uint64_t n, const TrainState* st, const float* w, float* ema,
                void* ema_half);

}  // namespace nlod
