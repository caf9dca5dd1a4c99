// Per-voxel visibility thresholds: for each voxel, the threshold t / (T - 1),
// t in [0, T), that maximises a classification score of the visibility
// network's probabilities against the ray-traced labels of that voxel's
// samples (rptr/cuda/neural-lod/neural_lod_learning.cu:122-266,
// compute_confusion_matrices / compute_scores / compute_optimal_thresholds;
// metrics neural_lod_learning_common.cuh:836-927).
//
// The reference adds one global atomic per (sample, threshold).  Here one
// workgroup owns a voxel: every sample adds 1 to an LDS histogram at
// k = #{t : p > fl(t / (T - 1))} (the thresholds it is predicted positive at,
// found by binary search on the same float thresholds), split by label; a scan
// of the two histograms gives TP(t) and FP(t) for every t at once
// (O(S + T) instead of O(S T)).  The scores and the plateau-centred argmax are
// the reference's float formulas and selection rule.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "nlod.h"
#include "nlod_common.h"

namespace nlod {
namespace {

constexpr int kThrThreads = 256;
constexpr uint32_t kMaxThresholds = 1024;

__device__ __forceinline__ float precision_score(uint32_t fp, uint32_t tp) {
    return tp == 0 ? 0.f : float(tp) / float(tp + fp);
}
__device__ __forceinline__ float recall_score(uint32_t fn, uint32_t tp) {
    return tp == 0 ? 0.f : float(tp) / float(tp + fn);
}
__device__ __forceinline__ float fscore(uint32_t tn, uint32_t fp, uint32_t fn, uint32_t tp, float beta) {
    const float p = precision_score(fp, tp), r = recall_score(fn, tp);
    if (p + r == 0.f) return 0.f;
    const float b2 = beta * beta;
    return ((1.f + b2) * p * r) / (b2 * p + r);
}

// neural_lod_learning_common.cuh:836-927, same operation order; uint32
// products wrap as there.
__device__ float metric_score(int metric, uint32_t tn, uint32_t fp, uint32_t fn, uint32_t tp, uint32_t sup_pos,
                              uint32_t sup_neg, float beta) {
    switch (metric) {
        case NLOD_METRIC_PRECISION: return precision_score(fp, tp);
        case NLOD_METRIC_RECALL: return recall_score(fn, tp);
        case NLOD_METRIC_FSCORE: return fscore(tn, fp, fn, tp, beta);
        case NLOD_METRIC_FSCORE_MACRO: return (fscore(tn, fp, fn, tp, beta) + fscore(tp, fn, fp, tn, beta)) * 0.5f;
        case NLOD_METRIC_FSCORE_WEIGHTED:
        case NLOD_METRIC_OPTIMAL_FSCORE_WEIGHTED: {
            const float fpos = fscore(tn, fp, fn, tp, beta);
            const float fneg = fscore(tp, fn, fp, tn, metric == NLOD_METRIC_FSCORE_WEIGHTED ? beta : 1.f);
            return (fpos * sup_pos + fneg * sup_neg) / (sup_pos + sup_neg);
        }
        case NLOD_METRIC_MCC: {
            const float numer = float(tp * tn) - float(fp * fn);
            const float den = float(tp + fp) * float(tp + fn) * float(tn + fp) * float(tn + fn);
            return den == 0 ? 0.f : numer / sqrtf(den);
        }
        case NLOD_METRIC_COHEN_KAPPA: {
            const float numer = 2.f * (float(tp * tn) - float(fn * fp));
            const float den = float(tp + fp) * float(fp + tn) + float(tp + fn) * float(fn + tn);
            return den == 0 ? 0.f : numer / den;
        }
        case NLOD_METRIC_ACCURACY: {
            const float numer = float(tp + tn), den = float(tp + tn + fp + fn);
            return den == 0 ? 0.f : numer / den;
        }
        case NLOD_METRIC_BALANCED_ACCURACY: {
            const float sens = tp == 0 ? 0.0f : float(tp) / float(tp + fn);
            const float spec = tn == 0 ? 0.0f : float(tn) / float(tn + fp);
            return (sens + spec) * 0.5f;
        }
        case NLOD_METRIC_FOWLKES_MALLOWS: return sqrtf(precision_score(fp, tp) * recall_score(fn, tp));
        default: {  // Youden's J
            const float sens = tp == 0 ? 0.0f : float(tp) / float(tp + fn);
            const float spec = tn == 0 ? 0.0f : float(tn) / float(tn + fp);
            return sens + spec - 1.f;
        }
    }
}

// One workgroup per voxel; samples of voxel v are labels / probas [v S, v S + S).
__global__ __launch_bounds__(kThrThreads) void thresholds_kernel(uint32_t S, uint32_t T, int metric, float beta,
                                                                 const float* __restrict__ labels,
                                                                 const float* __restrict__ probas,
                                                                 float* __restrict__ thresholds,
                                                                 float* __restrict__ scores_out) {
    __shared__ uint32_t hpos[kMaxThresholds + 1], hneg[kMaxThresholds + 1];
    __shared__ float score[kMaxThresholds];
    __shared__ uint32_t s_best;
    const uint32_t v = blockIdx.x;
    for (uint32_t t = threadIdx.x; t <= T; t += kThrThreads) hpos[t] = hneg[t] = 0;
    __syncthreads();
    const float steps = (float)(T - 1);
    for (uint32_t s = threadIdx.x; s < S; s += kThrThreads) {
        const size_t i = (size_t)v * S + s;
        const float p = probas[i];
        // k = number of t with p > t / (T - 1); the thresholds increase with t
        uint32_t lo = 0, hi = T;  // first t with !(p > thr(t)) lies in [lo, hi]
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (p > (float)mid / steps) lo = mid + 1;
            else hi = mid;
        }
        if (labels[i] == 1.0f) atomicAdd(&hpos[lo], 1u);
        else atomicAdd(&hneg[lo], 1u);
    }
    __syncthreads();
    // suffix sums: predicted positive at threshold t <=> k > t
    if (threadIdx.x == 0) {
        uint32_t ap = 0, an = 0;
        for (int k = (int)T; k >= 0; --k) {
            ap += hpos[k];
            an += hneg[k];
            hpos[k] = ap;  // now #{k' >= k}
            hneg[k] = an;
        }
    }
    __syncthreads();
    const uint32_t P = hpos[0], N = hneg[0];
    // entry e of the voxel's flat [T][TN, FP, FN, TP] confusion block
    auto cm = [&](uint32_t e) -> uint32_t {
        const uint32_t t = e / 4, c = e % 4;
        const uint32_t tp = hpos[t + 1], fp = hneg[t + 1];
        return c == 0 ? N - fp : c == 1 ? fp : c == 2 ? P - tp : tp;
    };
    // class supports as the reference reads them: entries (T - 1) * 2 .. + 3
    const uint32_t se = (T - 1) * 2;
    const uint32_t sup_pos = cm(se + 3) + cm(se + 2), sup_neg = cm(se + 1) + cm(se + 0);
    for (uint32_t t = threadIdx.x; t < T; t += kThrThreads) {
        const uint32_t tp = hpos[t + 1], fp = hneg[t + 1];
        const uint32_t fn = P - tp, tn = N - fp;
        const float sc = metric_score(metric, tn, fp, fn, tp, sup_pos, sup_neg, beta);
        score[t] = sc;
        if (scores_out) scores_out[(size_t)v * T + t] = sc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float best = -10.0f;  // every score is in [-1, 1]
        uint32_t opt = 0;
        for (uint32_t t = 0; t < T; ++t)
            if (score[t] > best) {
                best = score[t];
                opt = t;
            }
        uint32_t eq = 0;  // centre of the plateau of (nearly) equal scores that follows
        // the reference compares with the double literal 1e-4
        while (opt + eq < T && (double)fabsf(score[opt + eq] - score[opt]) < 1e-4) ++eq;
        s_best = opt + eq / 2;
    }
    __syncthreads();
    if (threadIdx.x == 0) thresholds[v] = (float)s_best / steps;
}

}  // namespace
}  // namespace nlod

extern "C" int nlod_optimal_thresholds(void* stream, uint32_t n_voxels, uint32_t samples_per_voxel,
                                       const float* d_labels, const float* d_probas, uint32_t n_thresholds,
                                       int metric, float beta, float* d_thresholds_out, float* d_scores_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(n_thresholds >= 2 && n_thresholds <= nlod::kMaxThresholds, "n_thresholds must be in [2, 1024]");
        NLOD_CHECK(metric >= NLOD_METRIC_PRECISION && metric <= NLOD_METRIC_YOUDEN_J, "unknown metric");
        if (n_voxels == 0) return;
        NLOD_CHECK(samples_per_voxel > 0, "samples_per_voxel must be positive");
        NLOD_CHECK(d_labels && d_probas && d_thresholds_out, "null buffer");
        hipLaunchKernelGGL(nlod::thresholds_kernel, dim3(n_voxels), dim3(nlod::kThrThreads), 0,
                           nlod::as_stream(stream), samples_per_voxel, n_thresholds, metric, beta, d_labels, d_probas,
                           d_thresholds_out, d_scores_out);
        NLOD_LAUNCH_CHECK();
    });
}
