/*
 * ORACLE — test infrastructure only.  Nothing in the product links, loads or
 * calls this file; only tests/ use it, and only as the checker.
 *
 * CPU restatement of the per-voxel threshold optimisation of the reference
 * (rptr/cuda/neural-lod/neural_lod_learning.cu:122-266): a [T][TN, FP, FN, TP]
 * confusion block per voxel counted sample by sample and threshold by
 * threshold (prediction p > t / (T - 1) as float), the metric of
 * neural_lod_learning_common.cuh:836-927 per threshold, then the first
 * maximum moved to the middle of the plateau of scores within 1e-4 after it.
 * Built with -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

static float precision_score(uint32_t fp, uint32_t tp) { return tp == 0 ? 0.f : (float)tp / (float)(tp + fp); }
static float recall_score(uint32_t fn, uint32_t tp) { return tp == 0 ? 0.f : (float)tp / (float)(tp + fn); }
static float fscore(uint32_t tn, uint32_t fp, uint32_t fn, uint32_t tp, float beta) {
    (void)tn;
    float p = precision_score(fp, tp), r = recall_score(fn, tp);
    if (p + r == 0.f) return 0.f;
    float b2 = beta * beta;
    return ((1.f + b2) * p * r) / (b2 * p + r);
}

static float score_of(int metric, uint32_t tn, uint32_t fp, uint32_t fn, uint32_t tp, uint32_t sp, uint32_t sn,
                      float beta) {
    switch (metric) {
        case 0: return precision_score(fp, tp);
        case 1: return recall_score(fn, tp);
        case 2: return fscore(tn, fp, fn, tp, beta);
        case 3: return (fscore(tn, fp, fn, tp, beta) + fscore(tp, fn, fp, tn, beta)) * 0.5f;
        case 4: return (fscore(tn, fp, fn, tp, beta) * sp + fscore(tp, fn, fp, tn, beta) * sn) / (sp + sn);
        case 5: return (fscore(tn, fp, fn, tp, beta) * sp + fscore(tp, fn, fp, tn, 1.f) * sn) / (sp + sn);
        case 6: {
            float numer = (float)(tp * tn) - (float)(fp * fn);
            float den = (float)(tp + fp) * (float)(tp + fn) * (float)(tn + fp) * (float)(tn + fn);
            return den == 0 ? 0.f : numer / sqrtf(den);
        }
        case 7: {
            float numer = 2.f * ((float)(tp * tn) - (float)(fn * fp));
            float den = (float)(tp + fp) * (float)(fp + tn) + (float)(tp + fn) * (float)(fn + tn);
            return den == 0 ? 0.f : numer / den;
        }
        case 8: {
            float numer = (float)(tp + tn), den = (float)(tp + tn + fp + fn);
            return den == 0 ? 0.f : numer / den;
        }
        case 9: {
            float se = tp == 0 ? 0.0f : (float)tp / (float)(tp + fn);
            float sp_ = tn == 0 ? 0.0f : (float)tn / (float)(tn + fp);
            return (se + sp_) * 0.5f;
        }
        case 10: return sqrtf(precision_score(fp, tp) * recall_score(fn, tp));
        default: {
            float se = tp == 0 ? 0.0f : (float)tp / (float)(tp + fn);
            float sp_ = tn == 0 ? 0.0f : (float)tn / (float)(tn + fp);
            return se + sp_ - 1.f;
        }
    }
}

void nlo_optimal_thresholds(uint32_t n_voxels, uint32_t S, const float* labels, const float* probas, uint32_t T,
                            int metric, float beta, float* thresholds, float* scores) {
    uint32_t* cm = (uint32_t*)malloc((size_t)T * 4 * sizeof(uint32_t));
    for (uint32_t v = 0; v < n_voxels; ++v) {
        for (uint32_t k = 0; k < T * 4; ++k) cm[k] = 0;
        for (uint32_t s = 0; s < S; ++s) {
            const size_t i = (size_t)v * S + s;
            const int truth = labels[i] == 1.0f;
            for (uint32_t t = 0; t < T; ++t) {
                const float threshold = (float)t / (float)(T - 1);
                const int pred = probas[i] > threshold;
                cm[t * 4 + (truth ? (pred ? 3 : 2) : (pred ? 1 : 0))]++;
            }
        }
        const uint32_t se = (T - 1) * 2;
        const uint32_t sp = cm[se + 3] + cm[se + 2], sn = cm[se + 1] + cm[se + 0];
        float* sc = scores + (size_t)v * T;
        for (uint32_t t = 0; t < T; ++t)
            sc[t] = score_of(metric, cm[t * 4], cm[t * 4 + 1], cm[t * 4 + 2], cm[t * 4 + 3], sp, sn, beta);
        float best = -10.0f;
        uint32_t opt = 0;
        for (uint32_t t = 0; t < T; ++t)
            if (sc[t] > best) { best = sc[t]; opt = t; }
        uint32_t eq = 0;
        while (opt + eq < T && fabsf(sc[opt + eq] - sc[opt]) < 1e-4) ++eq;
        thresholds[v] = (float)(opt + eq / 2) / (float)(T - 1);
    }
    free(cm);
}
