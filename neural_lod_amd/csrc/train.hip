// Training-step kernels: losses with FP64 loss reduction, the FP64 gradient
// norm / overflow check that drives dynamic loss scaling, Adam
// (tiny-cuda-nn include/tiny-cuda-nn/optimizers/adam.h:48-119) with the
// ExponentialDecay schedule (optimizers/exponential_decay.h:60-71) evaluated
// on the device, and the EMA of the weights (optimizers/ema.h:45-62).
//
// Every decision (skip the step, change the loss scale, the learning rate of
// this step) is taken on the device from TrainState, so a whole training step
// runs without a host round trip and can be captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "loss_device.h"
#include "train.h"

namespace nlod {
namespace {

constexpr int kBlock = 256;

template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / 64; ++w) s += red[w];
    __syncthreads();
    return s;  // valid in thread 0
}

// d_out = loss_scale * dL/dpred (float [n][dims]); per-block FP64 partial sums of the loss.
__global__ __launch_bounds__(kBlock) void loss_kernel(int type, uint64_t n_elems, uint32_t dims,
                                                      const TrainState* __restrict__ st, const float* __restrict__ pred,
                                                      const float* __restrict__ target, float* __restrict__ d_out,
                                                      double* __restrict__ partial) {
    __shared__ double red[kBlock / 64];
    const float scale = st->loss_scale;
    const float n_total = (float)n_elems;  // = N * dims
    double acc = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n_elems; i += (uint64_t)gridDim.x * kBlock) {
        double v;
        float g;
        loss_elem(type, pred[i], target[i], n_total, scale, v, g);
        d_out[i] = g;
        acc += v;
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
    (void)dims;
}

// Stand-alone loss evaluation (values and gradients per element).
__global__ __launch_bounds__(kBlock) void loss_eval_kernel(int type, uint64_t n_elems, float scale,
                                                           const float* __restrict__ pred,
                                                           const float* __restrict__ target, float* __restrict__ values,
                                                           float* __restrict__ grads) {
    const float n_total = (float)n_elems;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n_elems; i += (uint64_t)gridDim.x * kBlock) {
        double v;
        float g;
        loss_elem(type, pred[i], target[i], n_total, scale, v, g);
        if (values) values[i] = (float)v;
        grads[i] = g;
    }
}

// Per-block FP64 sum of g^2 and count of non-finite gradients.
__global__ __launch_bounds__(kBlock) void grad_stats_kernel(uint64_t n, const float* __restrict__ g,
                                                            double* __restrict__ partial_sq,
                                                            uint32_t* __restrict__ partial_bad) {
    __shared__ double red[kBlock / 64];
    __shared__ uint32_t redu[kBlock / 64];
    double acc = 0.0;
    uint32_t bad = 0;
    const uint64_t n4 = n / 4;
    const float4* g4 = reinterpret_cast<const float4*>(g);
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * kBlock) {
        const float4 v = g4[i];
        const float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (isfinite(e[k])) acc += (double)e[k] * (double)e[k];
            else ++bad;
        }
    }
    for (uint64_t i = n4 * 4 + (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        if (isfinite(g[i])) acc += (double)g[i] * (double)g[i];
        else ++bad;
    }
    const double s = block_sum(acc, red);
    const uint32_t b = block_sum(bad, redu);
    if (threadIdx.x == 0) {
        partial_sq[blockIdx.x] = s;
        partial_bad[blockIdx.x] = b;
    }
}

// One block: reduce the partials in a fixed order, decide skip / scale /
// learning rate for this step.
__global__ __launch_bounds__(kBlock) void finalize_kernel(TrainHyper h, uint32_t n_loss_partials,
                                                          const double* __restrict__ loss_partial,
                                                          uint32_t n_grad_partials,
                                                          const double* __restrict__ grad_partial,
                                                          const uint32_t* __restrict__ bad_partial,
                                                          TrainState* __restrict__ st) {
    __shared__ double red[kBlock / 64];
    __shared__ uint32_t redu[kBlock / 64];
    double l = 0.0, gq = 0.0;
    uint32_t bad = 0;
    // fixed per-thread strided order, then a fixed tree: deterministic
    for (uint32_t i = threadIdx.x; i < n_loss_partials; i += kBlock) l += loss_partial[i];
    for (uint32_t i = threadIdx.x; i < n_grad_partials; i += kBlock) {
        gq += grad_partial[i];
        bad += bad_partial[i];
    }
    l = block_sum(l, red);
    gq = block_sum(gq, red);
    bad = block_sum(bad, redu);
    if (threadIdx.x != 0) return;
    const float scale = st->loss_scale;
    st->applied_scale = scale;
    st->loss = l;
    const bool finite = bad == 0 && isfinite(gq) && isfinite(l);
    st->grad_sq = finite ? gq / ((double)scale * (double)scale) : (double)INFINITY;
    st->found_inf = finite ? 0u : 1u;
    if (!finite) {
        ++st->skipped;
        st->good_steps = 0;
        if (h.dynamic_scale) st->loss_scale = fmaxf(scale * h.backoff, h.min_scale);
        return;
    }
    // ExponentialDecay::step: the factor changes before the nested step, from the step count so far
    const uint32_t s = st->step;
    if (s == 0) st->lr_factor = 1.0f;
    if (h.decay_interval > 0 && s >= h.decay_start && (s - h.decay_start) % h.decay_interval == 0 &&
        s <= h.decay_end)
        st->lr_factor *= h.decay_base;
    st->lr = h.base_lr * st->lr_factor;
    st->step = s + 1;  // the optimizer's ++m_current_step
    if (h.dynamic_scale && ++st->good_steps >= h.growth_interval) {
        st->good_steps = 0;
        st->loss_scale = fminf(scale * h.growth, h.max_scale);
    }
}

// bias_correction(a, t): the step-size factor of adam.h's update for a
// parameter at its t-th step.  Parameters that took every step share t with
// the optimizer, so a kernel computes the factor for its step count T once and
// passes it as bc_T (the same float expression, hence the same value).
__device__ __forceinline__ float bias_correction(const AdamParams& a, uint32_t t) {
    return sqrtf(1.0f - powf(a.beta2, (float)t)) / (1.0f - powf(a.beta1, (float)t));
}

__device__ __forceinline__ float adam_one(const AdamParams& a, const TrainState& st, bool matrix, float w, float g,
                                          float& m, float& v, uint32_t& t, bool& skip, uint32_t T, float bc_T) {
    // adam.h:75-118
    float lr = st.lr;
    if (matrix) {
        if (!a.optimize_matrix) { skip = true; return w; }
        g += a.l2_reg * w;
    } else {
        if (!a.optimize_non_matrix || g == 0.0f) { skip = true; return w; }
        lr *= a.non_matrix_lr_factor;
    }
    skip = false;
    const float g2 = g * g;
    m = a.beta1 * m + (1.0f - a.beta1) * g;
    v = a.beta2 * v + (1.0f - a.beta2) * g2;
    t = t + 1;
    lr *= t == T ? bc_T : bias_correction(a, t);
    float lo = 0.0f, hi = 3.402823466e38f;
    if (a.adabound) {
        const float k = (1.0f - a.beta2) * (float)st.step;
        lo = 0.1f - 0.1f / (k + 1.0f);
        hi = 0.1f + 0.1f / k;
    }
    const float eff = fminf(fmaxf(lr / (sqrtf(v) + a.epsilon), lo), hi);
    const float decayed = (1.0f - a.relative_decay * lr) * w - copysignf(a.absolute_decay * lr, w);
    float nw = decayed - eff * m;
    if (a.clipping != 0.0f) nw = fminf(fmaxf(nw, -a.clipping), a.clipping);
    return nw;
}

// 4 parameters per thread.  CLEAR: zero the consumed gradient entries.
template <bool CLEAR>
__global__ __launch_bounds__(kBlock) void adam_kernel(AdamParams a, uint64_t n, uint64_t n_matrix,
                                                      const TrainState* __restrict__ stp, float* __restrict__ w,
                                                      _Float16* __restrict__ w_half, float* __restrict__ g,
                                                      float* __restrict__ m1, float* __restrict__ m2,
                                                      uint32_t* __restrict__ steps) {
    const TrainState st = *stp;
    const float scale = st.applied_scale;
    const uint64_t n4 = (n + 3) / 4;
    // the bias correction at the optimizer's step count, once per block
    __shared__ float s_bc;
    if (threadIdx.x == 0) s_bc = st.step ? bias_correction(a, st.step) : 0.0f;
    __syncthreads();
    const uint32_t T = st.step;
    const float bc_T = s_bc;
    for (uint64_t q = (uint64_t)blockIdx.x * kBlock + threadIdx.x; q < n4; q += (uint64_t)gridDim.x * kBlock) {
        const uint64_t i0 = q * 4;
        if (i0 + 4 <= n) {
            const float4 g4 = reinterpret_cast<const float4*>(g)[q];
            const float ge[4] = {g4.x, g4.y, g4.z, g4.w};
            const bool any = ge[0] != 0.0f || ge[1] != 0.0f || ge[2] != 0.0f || ge[3] != 0.0f;
            if (!any && i0 >= n_matrix) continue;  // untouched grid entries: nothing to read or write
            if (CLEAR && any) reinterpret_cast<float4*>(g)[q] = float4{0.f, 0.f, 0.f, 0.f};
            if (st.found_inf) continue;
            float4 w4 = reinterpret_cast<const float4*>(w)[q];
            float4 a4 = reinterpret_cast<const float4*>(m1)[q];
            float4 b4 = reinterpret_cast<const float4*>(m2)[q];
            uint4 t4 = reinterpret_cast<const uint4*>(steps)[q];
            float we[4] = {w4.x, w4.y, w4.z, w4.w}, me[4] = {a4.x, a4.y, a4.z, a4.w}, ve[4] = {b4.x, b4.y, b4.z, b4.w};
            uint32_t te[4] = {t4.x, t4.y, t4.z, t4.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                bool skip;
                float mk = me[k], vk = ve[k];
                uint32_t tk = te[k];
                const float nw = adam_one(a, st, i0 + k < n_matrix, we[k], ge[k] / scale, mk, vk, tk, skip, T, bc_T);
                if (!skip) {
                    we[k] = nw;
                    me[k] = mk;
                    ve[k] = vk;
                    te[k] = tk;
                }
            }
            reinterpret_cast<float4*>(w)[q] = float4{we[0], we[1], we[2], we[3]};
            reinterpret_cast<float4*>(m1)[q] = float4{me[0], me[1], me[2], me[3]};
            reinterpret_cast<float4*>(m2)[q] = float4{ve[0], ve[1], ve[2], ve[3]};
            reinterpret_cast<uint4*>(steps)[q] = uint4{te[0], te[1], te[2], te[3]};
            if (w_half) {
                typedef _Float16 h4 __attribute__((ext_vector_type(4)));
                reinterpret_cast<h4*>(w_half)[q] = h4{(_Float16)we[0], (_Float16)we[1], (_Float16)we[2], (_Float16)we[3]};
            }
        } else {
            for (uint64_t i = i0; i < n; ++i) {
                const float gi = g[i];
                if (gi == 0.0f && i >= n_matrix) continue;
                if (CLEAR && gi != 0.0f) g[i] = 0.0f;
                if (st.found_inf) continue;
                bool skip;
                float mk = m1[i], vk = m2[i];
                uint32_t tk = steps[i];
                const float nw = adam_one(a, st, i < n_matrix, w[i], gi / scale, mk, vk, tk, skip, T, bc_T);
                if (skip) continue;
                w[i] = nw;
                m1[i] = mk;
                m2[i] = vk;
                steps[i] = tk;
                if (w_half) w_half[i] = (_Float16)nw;
            }
        }
    }
}

// ema.h:45-62 (full precision): ema = (ema * decay * debias_old + w * (1 - decay)) * debias_new
__global__ __launch_bounds__(kBlock) void ema_kernel(float decay, uint64_t n, const TrainState* __restrict__ stp,
                                                     const float* __restrict__ w, float* __restrict__ ema,
                                                     _Float16* __restrict__ ema_half) {
    if (stp->found_inf) return;
    const uint32_t t = stp->step;
    const float debias_old = 1.0f - (float)pow((double)decay, (double)(t - 1));
    const float debias_new = 1.0f / (1.0f - (float)pow((double)decay, (double)t));
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        const float f = (ema[i] * decay * debias_old + w[i] * (1.0f - decay)) * debias_new;
        ema[i] = f;
        if (ema_half) ema_half[i] = (_Float16)f;
    }
}

__global__ void state_init_kernel(TrainState* st, float loss_scale) {
    TrainState s{};
    s.loss_scale = loss_scale;
    s.applied_scale = loss_scale;
    s.lr_factor = 1.0f;
    *st = s;
}

uint32_t blocks_for(uint64_t n_items, uint32_t cap) {
    const uint64_t b = (n_items + kBlock - 1) / kBlock;
    return (uint32_t)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

void train_state_init(hipStream_t s, TrainState* st, float loss_scale) {
    hipLaunchKernelGGL(state_init_kernel, dim3(1), dim3(1), 0, s, st, loss_scale);
    NLOD_LAUNCH_CHECK();
}

void loss_forward(hipStream_t s, int type, uint64_t n_elems, uint32_t dims, const TrainState* st, const float* pred,
                  const float* target, float* d_out, double* partial) {
    hipLaunchKernelGGL(loss_kernel, dim3(kLossPartials), dim3(kBlock), 0, s, type, n_elems, dims, st, pred, target,
                       d_out, partial);
    NLOD_LAUNCH_CHECK();
}

void loss_evaluate(hipStream_t s, int type, uint64_t n_elems, float scale, const float* pred, const float* target,
                   float* values, float* grads) {
    if (n_elems == 0) return;
    hipLaunchKernelGGL(loss_eval_kernel, dim3(blocks_for(n_elems, 4096)), dim3(kBlock), 0, s, type, n_elems, scale,
                       pred, target, values, grads);
    NLOD_LAUNCH_CHECK();
}

void grad_stats(hipStream_t s, uint64_t n, const float* g, double* partial_sq, uint32_t* partial_bad) {
    hipLaunchKernelGGL(grad_stats_kernel, dim3(kGradPartials), dim3(kBlock), 0, s, n, g, partial_sq, partial_bad);
    NLOD_LAUNCH_CHECK();
}

void train_finalize(hipStream_t s, const TrainHyper& h, uint32_t n_loss_partials, const double* loss_partial,
                    const double* grad_partial, const uint32_t* bad_partial, TrainState* st) {
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kBlock), 0, s, h, n_loss_partials, loss_partial,
                       (uint32_t)kGradPartials, grad_partial, bad_partial, st);
    NLOD_LAUNCH_CHECK();
}

void adam_update(hipStream_t s, const AdamParams& a, uint64_t n, uint64_t n_matrix, const TrainState* st, float* w,
                 void* w_half, float* g, float* m1, float* m2, uint32_t* steps, bool clear_grads) {
    if (n == 0) return;
    const uint32_t blocks = blocks_for((n + 3) / 4, 8192);
    auto* wh = static_cast<_Float16*>(w_half);
    if (clear_grads)
        hipLaunchKernelGGL(adam_kernel<true>, dim3(blocks), dim3(kBlock), 0, s, a, n, n_matrix, st, w, wh, g, m1, m2,
                           steps);
    else
        hipLaunchKernelGGL(adam_kernel<false>, dim3(blocks), dim3(kBlock), 0, s, a, n, n_matrix, st, w, wh, g, m1, m2,
                           steps);
    NLOD_LAUNCH_CHECK();
}

void ema_update(hipStream_t s, float decay, uint64_t n, const TrainState* st, const float* w, float* ema,
                void* ema_half) {
    if (n == 0) return;
    hipLaunchKernelGGL(ema_kernel, dim3(blocks_for(n, 8192)), dim3(kBlock), 0, s, decay, n, st, w, ema,
                       static_cast<_Float16*>(ema_half));
    NLOD_LAUNCH_CHECK();
}

}  // namespace nlod
