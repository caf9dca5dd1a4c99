// Fully fused FP16-MFMA training pass of the MLP (gfx950): forward, loss,
// backward and weight gradients of one batch in one kernel.  Activations and
// deltas never leave the CU; the weight gradients accumulate in MFMA
// accumulators over the block's queries.  Semantics: forward as
// tiny-cuda-nn fully_fused_mlp.cu:499-731, loss as the reference loss kernels,
// backward as fully_fused_mlp.cu:150-259,733-833 (no biases).
//
// Block = 8 waves (two per SIMD: while one waits on LDS, a barrier or a load,
// the other issues; each wave keeps 1/8 of the dW tiles, which keeps it within
// the 256 registers of two waves per SIMD); a round = 8 tiles of 16 queries.
// The next round's features and targets are loaded while the current round's
// backward runs.  Per round:
//   each wave: features X (B fragments) -> forward through the layers,
//   keeping every activation as the next layer's B fragment in registers ->
//   outputs -> loss value (FP64 partial) and dL/dy -> delta_out.
//   For each layer from the output down: the wave computes its delta from
//   the previous G and act'(H) in registers and stages delta and the layer's
//   input (H or X) in LDS unit-major [unit][64 queries]; after a barrier every
//   wave adds its share of the layer's dW tiles, dW += delta . input^T over the
//   round's 64 queries (v_mfma_f32_16x16x32_f16, queries as the reduction
//   dimension); then G = W^T delta for the next layer.
//   Finally dL/d(enc) = W0^T delta_0 goes to global memory (FP32 SoA) for the
//   encoding backward.
// At the end each block writes its dW sums to a partial buffer; a second
// kernel reduces the partials in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "loss_device.h"
#include "mfma_common.h"
#include "model.h"

namespace nlod {
namespace {

constexpr int kWaves = 8;  // two waves per SIMD
constexpr int kThreads = kWaves * 64;
constexpr int kQ = kWaves * 16;       // queries per round
constexpr int kStride = kQ + 8;       // LDS row stride in halves (16 B aligned rows)
constexpr int kMaxKTiles = 8;         // encoded width <= 128
constexpr uint32_t kMaxFusedBlocks = 512;

struct FusedArgs {
    uint32_t N, n_output, k_enc, k_padded;
    int32_t act, out_act, loss_type;
    uint32_t fwd_halves, bwd_halves;
    uint32_t n_net;
};

template <int W>
struct Shape {
    static constexpr int MT = W / 16;                        // 16-row tiles of a hidden layer
    static constexpr int KC = W / 32;                        // 32-wide k chunks of a hidden layer
    static constexpr int HT = (MT * MT + kWaves - 1) / kWaves;  // hidden dW tiles per wave
    static constexpr int ZT = (MT * kMaxKTiles + kWaves - 1) / kWaves;  // layer-0 dW tiles per wave (max)
};

// dW tile (mr, nc) += A[16 mr + m][q] * B[16 nc + n][q] over the round's queries.
__device__ __forceinline__ void dw_tile(const half_t* __restrict__ A, const half_t* __restrict__ B, int mr, int nc,
                                        floatx4& acc) {
    const int lane = threadIdx.x & 63, g = lane >> 4, m = lane & 15;
    const half_t* a = A + (size_t)(16 * mr + m) * kStride + 8 * g;
    const half_t* b = B + (size_t)(16 * nc + m) * kStride + 8 * g;
#pragma unroll
    for (int q0 = 0; q0 < kQ; q0 += 32) {
        const half8 av = *reinterpret_cast<const half8*>(a + q0);
        const half8 bv = *reinterpret_cast<const half8*>(b + q0);
        acc = mfma16(av, bv, acc);
    }
}

// Stages a B-fragment-layout activation (units 32c+4g+i, 32c+16+4g+i of query n) as rows.
template <int KC>
__device__ __forceinline__ void stage_units(half_t* __restrict__ buf, const half8* v, int col) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
#pragma unroll
    for (int c = 0; c < KC; ++c) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            buf[(size_t)(32 * c + 4 * g + i) * kStride + col] = v[c][i];
            buf[(size_t)(32 * c + 16 + 4 * g + i) * kStride + col] = v[c][i + 4];
        }
    }
}

// ReLU backward of one chunk on packed halves: the delta G (accumulator tiles
// lo, hi) rounded to FP16, times the step 1[h > 0] of the activation h.  ReLU
// outputs are +0, -0 or at least the smallest FP16 subnormal 2^-24, so
// min(h * 4096 * 4096, 1) is exactly 1 or a zero; the product with a zero
// step is a zero of either sign (only its sign can differ from rounding
// G * 0).  Packed FP16 arithmetic: three operations per two values for the
// step and one for the product.
__device__ __forceinline__ half8 relu_bwd_pack(const floatx4& lo, const floatx4& hi, const half8& h) {
    half8 d;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        d[i] = (half_t)lo[i];
        d[i + 4] = (half_t)hi[i];
    }
    const half8 big = {4096, 4096, 4096, 4096, 4096, 4096, 4096, 4096};
    const half8 one = {1, 1, 1, 1, 1, 1, 1, 1};
    const half8 step = __builtin_elementwise_min(h * big * big, one);
    return d * step;
}

// ACT: ACT_RELU fixed at compile time, or -1 to read the hidden activation
// from the arguments.
template <int W, int L, int ACT>
__global__ __launch_bounds__(kThreads) void fused_train_kernel(FusedArgs a, const half_t* __restrict__ packed,
                                                               const half_t* __restrict__ packed_bwd,
                                                               const half_t* __restrict__ feat,
                                                               const float* __restrict__ target,
                                                               const float* __restrict__ loss_scale,
                                                               double* __restrict__ loss_partial,
                                                               float* __restrict__ d_enc,
                                                               float* __restrict__ w_partial) {
    using S = Shape<W>;
    constexpr int MT = S::MT, KC = S::KC;
    extern __shared__ float4 smem4[];
    half_t* wf = reinterpret_cast<half_t*>(smem4);
    half_t* wb = wf + a.fwd_halves;
    half_t* bufA = wb + a.bwd_halves;                 // delta rows [W][kStride]
    half_t* bufB = bufA + (size_t)W * kStride;        // input rows [ROWS_B][kStride]
    __shared__ double red[kWaves];
    for (uint32_t k = threadIdx.x; k < a.fwd_halves / 8; k += kThreads)
        reinterpret_cast<float4*>(wf)[k] = reinterpret_cast<const float4*>(packed)[k];
    for (uint32_t k = threadIdx.x; k < a.bwd_halves / 8; k += kThreads)
        reinterpret_cast<float4*>(wb)[k] = reinterpret_cast<const float4*>(packed_bwd)[k];
    __syncthreads();

    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15, wave = threadIdx.x >> 6;
    const int act = ACT >= 0 ? ACT : a.act;
    const uint32_t kc0 = a.k_padded / 32, kp2 = a.k_padded / 2;
    const uint32_t KT = (a.k_enc + 15) / 16;  // layer-0 dW column tiles
    const uint32_t n_tiles = (a.N + 15) / 16;
    const uint32_t rounds = (n_tiles + gridDim.x * kWaves - 1) / (gridDim.x * kWaves);
    const float scale = *loss_scale;
    const float n_total = (float)((uint64_t)a.N * a.n_output);

    floatx4 acc_out = {0.f, 0.f, 0.f, 0.f};
    floatx4 acc_hid[L > 1 ? L - 1 : 1][S::HT];
    floatx4 acc_in[S::ZT];
#pragma unroll
    for (int l = 0; l < (L > 1 ? L - 1 : 1); ++l)
#pragma unroll
        for (int t = 0; t < S::HT; ++t) acc_hid[l][t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < S::ZT; ++t) acc_in[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    double lsum = 0.0;
    // single-output networks: only lane group 0 holds an output in a round, so
    // the loss values (the costly part of the loss) are evaluated every fourth
    // round with all four groups, one element each
    const bool single = a.n_output == 1;
    float pend_p = 0.0f, pend_t = 0.0f;
    bool pend_ok = false;

    // weight-fragment offsets (halves) of the forward and backward packings
    const size_t f_hidden0 = (size_t)MT * kc0 * 512;
    const size_t f_out = f_hidden0 + (size_t)(L - 1) * MT * KC * 512;
    const size_t b_hidden0 = (size_t)MT * 512;
    const size_t b_in = b_hidden0 + (size_t)(L - 1) * MT * KC * 512;

    // features (B fragments, k_padded <= 128: 4 chunks of 32) and targets of
    // a round's tile; the next round's are loaded during this round's backward
    half8 xb[4], xn[4];
    float tg[4], tn[4];
    auto load_round = [&](uint32_t r, half8(&x)[4], float(&t)[4]) {
        const uint32_t tile = (r * gridDim.x + blockIdx.x) * kWaves + wave;
        const bool ok = r < rounds && tile < n_tiles;
        const uint32_t q = tile * 16 + n;
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) x[c] = ok && c < kc0 ? load_feat_frag(feat, tile, kp2, c, g, n) : half8{};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t oi = 4 * g + i;
            t[i] = ok && q < a.N && oi < a.n_output ? target[(size_t)q * a.n_output + oi] : 0.0f;
        }
    };
    load_round(0, xb, tg);

    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t tile = (r * gridDim.x + blockIdx.x) * kWaves + wave;
        const bool tile_ok = tile < n_tiles;
        const uint32_t q = tile * 16 + n;
        const bool valid = tile_ok && q < a.N;
        const int col = wave * 16 + n;

        // ---- forward
        floatx4 acc[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
            if (c >= kc0) break;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const half8 w = *reinterpret_cast<const half8*>(wf + ((size_t)(mt * kc0 + c) * 64 + lane) * 8);
                acc[mt] = mfma16(w, xb[c], acc[mt]);
            }
        }
        half8 h[L][KC];
#pragma unroll
        for (int c = 0; c < KC; ++c) h[0][c] = pack_act(acc[2 * c], acc[2 * c + 1], act);
#pragma unroll
        for (int l = 1; l < L; ++l) {
            const half_t* fr = wf + f_hidden0 + (size_t)(l - 1) * MT * KC * 512;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                floatx4 s = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int c = 0; c < KC; ++c) {
                    const half8 w = *reinterpret_cast<const half8*>(fr + ((size_t)(mt * KC + c) * 64 + lane) * 8);
                    s = mfma16(w, h[l - 1][c], s);
                }
                acc[mt] = s;
            }
#pragma unroll
            for (int c = 0; c < KC; ++c) h[l][c] = pack_act(acc[2 * c], acc[2 * c + 1], act);
        }
        floatx4 o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            const half8 w = *reinterpret_cast<const half8*>(wf + f_out + ((size_t)c * 64 + lane) * 8);
            o = mfma16(w, h[L - 1][c], o);
        }

        // ---- loss and delta_out (rows 4g+i = outputs of query n)
        half8 bo;
        float p0 = 0.0f;  // output 0 of query n after the output activation (group 0)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t oi = 4 * g + i;
            float d = 0.0f;
            if (valid && oi < a.n_output) {
                const float p = act_fwd(a.out_act, o[i]);
                const float gr = loss_grad(a.loss_type, p, tg[i], n_total, scale);
                if (single) {
                    p0 = p;
                } else {
                    lsum += loss_value(a.loss_type, p, tg[i], n_total);
                }
                d = gr * act_bwd(a.out_act, p);
            }
            bo[i] = (half_t)d;
            bo[i + 4] = (half_t)0.0f;
        }
        if (single) {
            // lane (g, n) keeps the element of round 4k + g and evaluates the
            // values of four rounds at once
            const uint32_t slot = r & 3u;
            const float sp = __shfl(p0, n, 64), st = __shfl(tg[0], n, 64);
            if ((uint32_t)g == slot) {
                pend_p = sp;
                pend_t = st;
                pend_ok = valid;
            }
            if (slot == 3u || r + 1 == rounds) {
                if (pend_ok) lsum += loss_value(a.loss_type, pend_p, pend_t, n_total);
                pend_ok = false;
            }
        }

        load_round(r + 1, xn, tn);

        // ---- output layer: dW_out = delta_out . H_{L-1}^T
#pragma unroll
        for (int i = 0; i < 4; ++i) bufA[(size_t)(4 * g + i) * kStride + col] = bo[i];
        stage_units<KC>(bufB, h[L - 1], col);
        __syncthreads();
        if (wave < MT) dw_tile(bufA, bufB, 0, wave, acc_out);
        __syncthreads();
        // G = W_out^T delta_out
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const half8 w = *reinterpret_cast<const half8*>(wb + ((size_t)mt * 64 + lane) * 8);
            acc[mt] = mfma16(w, bo, floatx4{0.f, 0.f, 0.f, 0.f});
        }

        // ---- hidden layers L-1 .. 0
        half8 dl[KC];
#pragma unroll
        for (int l = L - 1; l >= 0; --l) {
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                if constexpr (ACT == ACT_RELU) {
                    dl[c] = relu_bwd_pack(acc[2 * c], acc[2 * c + 1], h[l][c]);
                } else {
                    float y[8], d[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) y[i] = (float)h[l][c][i];
                    act_bwd_n<8>(act, y, d);
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        dl[c][i] = (half_t)(acc[2 * c][i] * d[i]);
                        dl[c][i + 4] = (half_t)(acc[2 * c + 1][i] * d[i + 4]);
                    }
                }
            }
            stage_units<KC>(bufA, dl, col);
            if (l > 0) {
                stage_units<KC>(bufB, h[l - 1], col);
            } else {
                // layer 0 input: the features, natural order (feature 32c + 8g + j of query n)
#pragma unroll
                for (uint32_t c = 0; c < 4; ++c) {
                    if (c >= kc0) break;
#pragma unroll
                    for (int j = 0; j < 8; ++j) bufB[(size_t)(32 * c + 8 * g + j) * kStride + col] = xb[c][j];
                }
            }
            __syncthreads();
            if (l > 0) {
#pragma unroll
                for (int t = 0; t < S::HT; ++t) {
                    const int ti = wave + kWaves * t;
                    if (ti < MT * MT) dw_tile(bufA, bufB, ti / MT, ti % MT, acc_hid[l - 1][t]);
                }
            } else {
#pragma unroll
                for (int t = 0; t < S::ZT; ++t) {
                    const int ti = wave + kWaves * t;
                    if (ti < (int)(MT * KT)) dw_tile(bufA, bufB, ti / (int)KT, ti % (int)KT, acc_in[t]);
                }
            }
            __syncthreads();
            if (l > 0) {
                // G = W_l^T delta_l
                const half_t* fr = wb + b_hidden0 + (size_t)(L - 1 - l) * MT * KC * 512;
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    floatx4 s = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int c = 0; c < KC; ++c) {
                        const half8 w = *reinterpret_cast<const half8*>(fr + ((size_t)(mt * KC + c) * 64 + lane) * 8);
                        s = mfma16(w, dl[c], s);
                    }
                    acc[mt] = s;
                }
            }
        }

        // ---- dL/d(enc) = W0^T delta_0, rows = encoded features
        if (d_enc) {
            const uint32_t mt0 = a.k_padded / 16;
            for (uint32_t mt = 0; mt < mt0; ++mt) {
                floatx4 s = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int c = 0; c < KC; ++c) {
                    const half8 w = *reinterpret_cast<const half8*>(wb + b_in + ((size_t)(mt * KC + c) * 64 + lane) * 8);
                    s = mfma16(w, dl[c], s);
                }
                if (valid) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const uint32_t row = 16 * mt + 4 * g + i;
                        if (row < a.k_enc) d_enc[(size_t)row * a.N + q] = s[i];
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) xb[c] = xn[c];
#pragma unroll
        for (int i = 0; i < 4; ++i) tg[i] = tn[i];
    }

    // ---- per-block results: loss partial and dW partial (row-major matrices)
#pragma unroll
    for (int o2 = 32; o2 > 0; o2 >>= 1) lsum += __shfl_down(lsum, o2, 64);
    if (lane == 0) red[wave] = lsum;
    float* wp = w_partial + (size_t)blockIdx.x * a.n_net;
    const size_t off_hidden = (size_t)W * a.k_enc;
    const size_t off_out = off_hidden + (size_t)(L - 1) * W * W;
    // output layer: rows = outputs (16 padded), cols = W
    if (wave < MT) {
#pragma unroll
        for (int i = 0; i < 4; ++i) wp[off_out + (size_t)(4 * g + i) * W + 16 * wave + n] = acc_out[i];
    }
#pragma unroll
    for (int l = 1; l < L; ++l) {
#pragma unroll
        for (int t = 0; t < S::HT; ++t) {
            const int ti = wave + kWaves * t;
            if (ti >= MT * MT) continue;
            const int mr = ti / MT, nc = ti % MT;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                wp[off_hidden + (size_t)(l - 1) * W * W + (size_t)(16 * mr + 4 * g + i) * W + 16 * nc + n] =
                    acc_hid[l - 1][t][i];
        }
    }
#pragma unroll
    for (int t = 0; t < S::ZT; ++t) {
        const int ti = wave + kWaves * t;
        if (ti >= (int)(MT * KT)) continue;
        const int mr = ti / (int)KT, nc = ti % (int)KT;
        const uint32_t colk = 16 * nc + n;
        if (colk < a.k_enc) {
#pragma unroll
            for (int i = 0; i < 4; ++i) wp[(size_t)(16 * mr + 4 * g + i) * a.k_enc + colk] = acc_in[t][i];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kWaves; ++w) s += red[w];
        loss_partial[blockIdx.x] = s;
    }
}

// out[e] (+)= sum of the per-block weight-gradient rows.  A block takes 64
// consecutive parameters; its 16 row slices (the four 16-lane groups of each
// of its four waves) each sum a contiguous range of rows, four parameters per
// lane and 8 rows in flight, and the slices are added in a fixed order, so the
// result does not depend on timing.
constexpr int kReduceSlices = 16;
constexpr int kReduceCols = 64;

__global__ __launch_bounds__(256) void partial_reduce_kernel(uint32_t n, uint32_t rows,
                                                             const float* __restrict__ partial,
                                                             float* __restrict__ out, int accumulate) {
    __shared__ float4 s_sum[kReduceSlices][kReduceCols / 4];
    const uint32_t slice = threadIdx.x / 16, col4 = threadIdx.x % 16;
    const uint32_t e0 = blockIdx.x * kReduceCols + col4 * 4;
    const uint32_t per = (rows + kReduceSlices - 1) / kReduceSlices;
    const uint32_t r0 = min(rows, slice * per), r1 = min(rows, r0 + per);
    float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    if ((n & 3) == 0 && e0 + 4 <= n) {
        const float4* p = reinterpret_cast<const float4*>(partial + e0);
        const size_t stride4 = n / 4;
        uint32_t r = r0;
        for (; r + 8 <= r1; r += 8) {
            float4 v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = p[(size_t)(r + k) * stride4];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                acc.x += v[k].x;
                acc.y += v[k].y;
                acc.z += v[k].z;
                acc.w += v[k].w;
            }
        }
        for (; r < r1; ++r) {
            const float4 v = p[(size_t)r * stride4];
            acc.x += v.x;
            acc.y += v.y;
            acc.z += v.z;
            acc.w += v.w;
        }
    } else {
        for (uint32_t r = r0; r < r1; ++r) {
            const float* row = partial + (size_t)r * n;
            if (e0 + 0 < n) acc.x += row[e0 + 0];
            if (e0 + 1 < n) acc.y += row[e0 + 1];
            if (e0 + 2 < n) acc.z += row[e0 + 2];
            if (e0 + 3 < n) acc.w += row[e0 + 3];
        }
    }
    s_sum[slice][col4] = acc;
    __syncthreads();
    if (threadIdx.x >= kReduceCols) return;
    const uint32_t e = blockIdx.x * kReduceCols + threadIdx.x;
    if (e >= n) return;
    const float* flat = reinterpret_cast<const float*>(&s_sum[0][0]);
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < kReduceSlices; ++k) sum += flat[k * kReduceCols + threadIdx.x];
    out[e] = accumulate ? out[e] + sum : sum;
}

template <typename Fn>
void dispatch_wl(uint32_t W, uint32_t L, Fn&& f) {
    auto with_l = [&](auto Wc) {
        switch (L) {
            case 1: f(Wc, std::integral_constant<int, 1>{}); break;
            case 2: f(Wc, std::integral_constant<int, 2>{}); break;
            case 3: f(Wc, std::integral_constant<int, 3>{}); break;
            case 4: f(Wc, std::integral_constant<int, 4>{}); break;
            case 5: f(Wc, std::integral_constant<int, 5>{}); break;
            case 6: f(Wc, std::integral_constant<int, 6>{}); break;
            case 7: f(Wc, std::integral_constant<int, 7>{}); break;
            case 8: f(Wc, std::integral_constant<int, 8>{}); break;
            default: fail(NLOD_ERR_UNSUPPORTED, "fused training supports 1-8 hidden layers");
        }
    };
    switch (W) {
        case 32: with_l(std::integral_constant<int, 32>{}); break;
        case 64: with_l(std::integral_constant<int, 64>{}); break;
        default: fail(NLOD_ERR_UNSUPPORTED, "fused training supports widths 32 and 64");
    }
}

size_t fused_smem(const Layout& L) {
    const uint32_t W = L.net.width;
    const uint32_t rows_b = std::max<uint32_t>(W, k_padded_of(L.net));
    return mfma_packed_bytes(L.net) + mfma_packed_bwd_bytes(L.net) + sizeof(half_t) * (size_t)(W + rows_b) * kStride;
}

}  // namespace

bool mfma_fused_train_supported(const Layout& L) {
    const NetShape& n = L.net;
    if (n.width != 32 && n.width != 64) return false;
    // register budget of one wave: activations of every layer plus its dW tiles
    if (n.n_hidden_layers < 1 || n.n_hidden_layers > (n.width == 64 ? 5u : 8u)) return false;
    if (k_padded_of(n) > 16 * kMaxKTiles) return false;
    return fused_smem(L) <= 160 * 1024;
}

uint32_t mfma_fused_train_blocks(const Layout& L, uint32_t N) {
    const uint32_t rounds_min = div_round_up(div_round_up(std::max<uint32_t>(N, 1), 16), kWaves);
    return std::max<uint32_t>(1, std::min<uint32_t>({rounds_min, (uint32_t)cu_count(), kMaxFusedBlocks}));
}

void mfma_fused_train(hipStream_t stream, const Layout& L, const void* packed, const void* packed_bwd,
                      const void* feat, const float* target, uint32_t N, int loss_type, const float* loss_scale,
                      double* loss_partial, float* d_enc, float* w_partial) {
    if (!mfma_fused_train_supported(L)) fail(NLOD_ERR_UNSUPPORTED, "layout not supported by the fused training kernel");
    FusedArgs a{};
    a.N = N;
    a.n_output = L.net.n_output;
    a.k_enc = L.net.n_input;
    a.k_padded = k_padded_of(L.net);
    a.act = L.net.activation;
    a.out_act = L.net.output_activation;
    a.loss_type = loss_type;
    a.fwd_halves = (uint32_t)(mfma_packed_bytes(L.net) / sizeof(half_t));
    a.bwd_halves = (uint32_t)(mfma_packed_bwd_bytes(L.net) / sizeof(half_t));
    a.n_net = (uint32_t)L.net.n_params();
    const size_t smem = fused_smem(L);
    const uint32_t blocks = mfma_fused_train_blocks(L, N);
    dispatch_wl(L.net.width, L.net.n_hidden_layers, [&](auto Wc, auto Lc) {
        constexpr int W = decltype(Wc)::value, LL = decltype(Lc)::value;
        if constexpr (W == 64 && LL > 5) {
            fail(NLOD_ERR_UNSUPPORTED, "fused training supports at most 5 hidden layers at width 64");
        } else {
            auto kern = a.act == ACT_RELU ? fused_train_kernel<W, LL, ACT_RELU> : fused_train_kernel<W, LL, -1>;
            (void)occupancy_blocks((const void*)kern, kThreads, smem);  // raises the dynamic LDS limit
            hipLaunchKernelGGL(kern, dim3(blocks), dim3(kThreads), smem, stream, a,
                               static_cast<const half_t*>(packed), static_cast<const half_t*>(packed_bwd),
                               static_cast<const half_t*>(feat), target, loss_scale, loss_partial, d_enc, w_partial);
        }
    });
    NLOD_LAUNCH_CHECK();
}

void partial_reduce(hipStream_t stream, uint32_t n, uint32_t rows, const float* partial, float* out, bool accumulate) {
    if (n == 0) return;
    hipLaunchKernelGGL(partial_reduce_kernel, dim3(div_round_up(n, kReduceCols)), dim3(256), 0, stream, n, rows,
                       partial, out, accumulate ? 1 : 0);
    NLOD_LAUNCH_CHECK();
}

}  // namespace nlod
