// Device helpers shared by the encoding and MLP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.h"

namespace nlod {

typedef _Float16 half_t;
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// One grid segment as seen by a kernel.
struct GridArgs {
    GridLevels lv;
    uint32_t in_offset;   // first input column of the 3 positions
    uint32_t in_stride;   // floats per input row (AoS)
    uint32_t out_offset;  // first encoded row
    uint32_t pad_;
};

// pos_fract (tiny-cuda-nn common_device.h:438-450): pos = x*scale + 0.5 as one
// fused multiply-add, cell = floor(pos), fraction in float; smoothstep optional.
__device__ __forceinline__ void grid_pos(float x, float scale, int interp, float& frac, uint32_t& cell) {
#pragma clang fp contract(off)
    const float p = __builtin_fmaf(x, scale, 0.5f);
    const float fl = floorf(p);
    cell = (uint32_t)(int)fl;
    float f = p - fl;
    if (interp == INTERP_SMOOTHSTEP) f = f * f * (3.0f - 2.0f * f);
    frac = f;
}

// grid_index (tiny-cuda-nn encodings/grid.h:170-187) with the coherent prime
// hash {1, 2654435761, 805459861}; hs = entries of the level.
__device__ __forceinline__ uint32_t grid_index(int grid_type, uint32_t hs, uint32_t res, uint32_t x, uint32_t y,
                                               uint32_t z) {
    // the reference's stride loop, unrolled, with the same uint32 wrap-around
    uint32_t index = x;
    uint32_t stride = res;
    if (stride <= hs) {
        index += y * stride;
        stride *= res;
        if (stride <= hs) {
            index += z * stride;
            stride *= res;
        }
    }
    if (grid_type == GRID_HASH && hs < stride) index = x ^ (y * 2654435761u) ^ (z * 805459861u);
    // hashed levels have power-of-two size; dense ones only wrap at the far corner
    if ((hs & (hs - 1)) == 0) return index & (hs - 1);
    return index < hs ? index : index % hs;
}

// grid_index for the corners of one cell, with the per-level part done once:
// the stride loop's outcome (which strides it added) and the choice of the
// hash depend only on the level, and (y + 1) * m = y * m + m in uint32
// arithmetic, so a cell's rows cost adds instead of multiplies.  The indices
// equal grid_index's for every input.
struct LevelIndexer {
    uint32_t hs, my, mz;  // hashed: the primes; dense: the strides the loop added (0 if not)
    bool hash;
};

__device__ __forceinline__ LevelIndexer level_indexer(int grid_type, uint32_t hs, uint32_t res) {
    uint32_t s1 = 0, s2 = 0, stride = res;
    if (stride <= hs) {
        s1 = stride;
        stride *= res;
        if (stride <= hs) {
            s2 = stride;
            stride *= res;
        }
    }
    LevelIndexer li;
    li.hs = hs;
    li.hash = grid_type == GRID_HASH && hs < stride;
    li.my = li.hash ? 2654435761u : s1;
    li.mz = li.hash ? 805459861u : s2;
    return li;
}

// Index of corner x of the row whose y and z terms are ym = y * my, zm = z * mz.
__device__ __forceinline__ uint32_t level_index(const LevelIndexer& li, uint32_t x, uint32_t ym, uint32_t zm) {
    uint32_t index = li.hash ? (x ^ ym ^ zm) : (x + ym + zm);
    if ((li.hs & (li.hs - 1)) == 0) return index & (li.hs - 1);
    // dense levels wrap only for positions outside the unit cube: a branch,
    // so in-range queries never run the division
    if (__builtin_expect(index >= li.hs, 0)) index %= li.hs;
    return index;
}

__device__ __forceinline__ float act_fwd(int a, float v) {
    switch (a) {
        case ACT_RELU: return v > 0.0f ? v : 0.0f;
        case ACT_SIGMOID: return 1.0f / (1.0f + __expf(-v));
        case ACT_EXP: return __expf(v);
        default: return v;
    }
}

// act_fwd over N values, the switch taken once for all of them: the activation
// is uniform over the wave, and per element the compiler branches around every
// value, which multiplies the instruction count of an MFMA layer several times.
template <int N>
__device__ __forceinline__ void act_fwd_n(int a, float* v) {
    switch (a) {
        case ACT_RELU:
#pragma unroll
            for (int i = 0; i < N; ++i) v[i] = v[i] > 0.0f ? v[i] : 0.0f;
            break;
        case ACT_SIGMOID:
#pragma unroll
            for (int i = 0; i < N; ++i) v[i] = 1.0f / (1.0f + __expf(-v[i]));
            break;
        case ACT_EXP:
#pragma unroll
            for (int i = 0; i < N; ++i) v[i] = __expf(v[i]);
            break;
        default: break;
    }
}

// derivative from the post-activation value y (common_device.h:193-244)
__device__ __forceinline__ float act_bwd(int a, float y) {
    switch (a) {
        case ACT_RELU: return y > 0.0f ? 1.0f : 0.0f;
        case ACT_SIGMOID: return y * (1.0f - y);
        case ACT_EXP: return y;
        default: return 1.0f;
    }
}

// act_bwd over N values (d[i] = act'(y[i])), the switch taken once
template <int N>
__device__ __forceinline__ void act_bwd_n(int a, const float* y, float* d) {
    switch (a) {
        case ACT_RELU:
#pragma unroll
            for (int i = 0; i < N; ++i) d[i] = y[i] > 0.0f ? 1.0f : 0.0f;
            break;
        case ACT_SIGMOID:
#pragma unroll
            for (int i = 0; i < N; ++i) d[i] = y[i] * (1.0f - y[i]);
            break;
        case ACT_EXP:
#pragma unroll
            for (int i = 0; i < N; ++i) d[i] = y[i];
            break;
        default:
#pragma unroll
            for (int i = 0; i < N; ++i) d[i] = 1.0f;
            break;
    }
}

}  // namespace nlod
