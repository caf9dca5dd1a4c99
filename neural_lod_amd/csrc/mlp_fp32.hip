// Plain FP32 form of the fully fused MLP (tiny-cuda-nn src/fully_fused_mlp.cu,
// no biases): forward, backward through the layers, and weight gradients.
//
// A workgroup of 256 lanes owns a tile of T queries and keeps the tile's
// activations in LDS ([unit][query], FP32).  Each layer is an LDS-resident
// GEMM out[W][T] = A[W][K] . in[K][T]: the current weight matrix is staged
// k-major in LDS, and every lane accumulates a (W/16 units x T/16 queries)
// register block with FP32 FMAs (one float4 of weights and T/64 float4s of
// activations per 4*(T/16)... FMAs).  Weight gradients are batch reductions
// dW = delta . act^T computed as split-N tiled FP32 GEMMs with a
// deterministic second-pass reduction.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "kernels.h"
#include "nlod_common.h"

namespace nlod {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kChunkK = 64;  // layer-0 / d_enc rows handled per pass

struct MlpArgs {
    uint32_t n_input, n_hidden_layers, n_output, n_output_padded;
    int32_t act, out_act;
};

template <int W>
struct Tiling {
    static constexpr int UB = W / 16;             // units per lane
    static constexpr int QB = W >= 128 ? 4 : 8;  // queries per lane
    static constexpr int T = 16 * QB;            // queries per tile
};

// acc[UB][QB] += sum_k AT[k][u0 + i] * X[k][q0 + j]; AT is [kk][W] (k-major), X is [kk][T]
template <int W>
__device__ __forceinline__ void lds_gemm(const float* __restrict__ AT, const float* __restrict__ X, uint32_t kk,
                                         int u0, int q0, float (&acc)[Tiling<W>::UB][Tiling<W>::QB]) {
    constexpr int UB = Tiling<W>::UB, QB = Tiling<W>::QB, T = Tiling<W>::T;
#pragma unroll 2
    for (uint32_t k = 0; k < kk; ++k) {
        float a[UB], x[QB];
#pragma unroll
        for (int i = 0; i < UB; i += (UB >= 4 ? 4 : UB)) {
            if constexpr (UB >= 4) {
                const float4 v = *reinterpret_cast<const float4*>(AT + k * W + u0 + i);
                a[i] = v.x; a[i + 1] = v.y; a[i + 2] = v.z; a[i + 3] = v.w;
            } else {
#pragma unroll
                for (int t = 0; t < UB; ++t) a[t] = AT[k * W + u0 + t];
            }
        }
#pragma unroll
        for (int j = 0; j < QB; j += 4) {
            const float4 v = *reinterpret_cast<const float4*>(X + k * T + q0 + j);
            x[j] = v.x; x[j + 1] = v.y; x[j + 2] = v.z; x[j + 3] = v.w;
        }
#pragma unroll
        for (int i = 0; i < UB; ++i)
#pragma unroll
            for (int j = 0; j < QB; ++j) acc[i][j] += a[i] * x[j];
    }
}

// Stage rows [r0, r0+rows) x cols of a row-major matrix (stride ld) into LDS,
// transposed to [col][W-strided row]: dst[c * W + (r - r0)].
__device__ __forceinline__ void stage_transposed(float* dst, const float* __restrict__ src, uint32_t rows,
                                                 uint32_t cols, uint32_t ld, uint32_t col0, uint32_t Wd) {
    for (uint32_t idx = threadIdx.x; idx < rows * cols; idx += kThreads) {
        const uint32_t r = idx / cols, c = idx - r * cols;
        dst[c * Wd + r] = src[(size_t)r * ld + col0 + c];
    }
}

// Stage rows [k0, k0+kk) of a row-major [*, ld] matrix as-is: dst[k * ld + m].
__device__ __forceinline__ void stage_rows(float* dst, const float* __restrict__ src, uint32_t kk, uint32_t ld,
                                           uint32_t m0, uint32_t mm, uint32_t Wd) {
    for (uint32_t idx = threadIdx.x; idx < kk * mm; idx += kThreads) {
        const uint32_t k = idx / mm, m = idx - k * mm;
        dst[k * Wd + m] = src[(size_t)k * ld + m0 + m];
    }
}

// Load rows [k0, k0+kk) of a SoA [*][N] global matrix for the tile into X[kk][T] (0 beyond N).
template <int T>
__device__ __forceinline__ void load_tile_rows(float* X, const float* __restrict__ src, uint32_t k0, uint32_t kk,
                                               uint32_t N, uint32_t tq0) {
    for (uint32_t idx = threadIdx.x; idx < kk * T; idx += kThreads) {
        const uint32_t k = idx / T, q = idx - k * T;
        X[idx] = (tq0 + q < N) ? src[(size_t)(k0 + k) * N + tq0 + q] : 0.0f;
    }
}

template <int W>
__global__ __launch_bounds__(kThreads) void mlp_fwd_kernel(uint32_t N, MlpArgs a, const float* __restrict__ w,
                                                           const float* __restrict__ enc, float* __restrict__ out,
                                                           float* __restrict__ hidden) {
    using Tl = Tiling<W>;
    constexpr int UB = Tl::UB, QB = Tl::QB, T = Tl::T;
    constexpr int R = W > (int)kChunkK ? W : (int)kChunkK;  // rows of the staging buffers
    extern __shared__ float4 smem4[];
    float* bufA = reinterpret_cast<float*>(smem4);  // [R][T]
    float* bufB = bufA + R * T;                      // [W][T]
    float* wt = bufB + W * T;                        // [R][W]
    const int tu = threadIdx.x >> 4, tq = threadIdx.x & 15;
    const int u0 = tu * UB, q0 = tq * QB;
    const uint32_t K = a.n_input;
    const uint32_t n_tiles = (N + T - 1) / T;
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint32_t tq0 = tile * T;
        float acc[UB][QB];
#pragma unroll
        for (int i = 0; i < UB; ++i)
#pragma unroll
            for (int j = 0; j < QB; ++j) acc[i][j] = 0.0f;
        for (uint32_t k0 = 0; k0 < K; k0 += kChunkK) {
            const uint32_t kc = min(kChunkK, K - k0);
            __syncthreads();
            load_tile_rows<T>(bufA, enc, k0, kc, N, tq0);
            stage_transposed(wt, w, W, kc, K, k0, W);
            __syncthreads();
            lds_gemm<W>(wt, bufA, kc, u0, q0, acc);
        }
        float* cur = bufB;
        float* nxt = bufA;
        for (uint32_t layer = 0;; ++layer) {
            // activation -> LDS tile (and the saved activations for backward)
#pragma unroll
            for (int i = 0; i < UB; ++i) act_fwd_n<QB>(a.act, acc[i]);
#pragma unroll
            for (int i = 0; i < UB; ++i)
#pragma unroll
                for (int j = 0; j < QB; ++j) {
                    const float h = acc[i][j];
                    cur[(u0 + i) * T + q0 + j] = h;
                    if (hidden && tq0 + q0 + j < N) hidden[((size_t)layer * W + u0 + i) * N + tq0 + q0 + j] = h;
                }
            if (layer + 1 >= a.n_hidden_layers) break;
            __syncthreads();
            stage_transposed(wt, w + (size_t)W * K + (size_t)layer * W * W, W, W, W, 0, W);
            __syncthreads();
#pragma unroll
            for (int i = 0; i < UB; ++i)
#pragma unroll
                for (int j = 0; j < QB; ++j) acc[i][j] = 0.0f;
            lds_gemm<W>(wt, cur, W, u0, q0, acc);
            float* t = cur;
            cur = nxt;
            nxt = t;
        }
        __syncthreads();
        // output layer: lane t < T computes query t's outputs (weights broadcast from global/L1)
        const float* w_out = w + (size_t)W * K + (size_t)(a.n_hidden_layers - 1) * W * W;
        for (uint32_t t = threadIdx.x; t < (uint32_t)T; t += kThreads) {
            if (tq0 + t >= N) continue;
            for (uint32_t r = 0; r < a.n_output; ++r) {
                float s = 0.0f;
#pragma unroll 8
                for (int k = 0; k < W; ++k) s += w_out[r * W + k] * cur[k * T + t];
                out[(size_t)(tq0 + t) * a.n_output + r] = act_fwd(a.out_act, s);
            }
        }
    }
}

// Backward through the layers for a tile: delta_l (SoA [L][W][N], dL/d pre-activation),
// delta_out (SoA [n_out_padded][N]) and dL/d(enc) (SoA [K][N], optional).
template <int W>
__global__ __launch_bounds__(kThreads) void mlp_bwd_delta_kernel(uint32_t N, MlpArgs a, const float* __restrict__ w,
                                                                 const float* __restrict__ hidden,
                                                                 const float* __restrict__ out,
                                                                 const float* __restrict__ d_out,
                                                                 float* __restrict__ delta,
                                                                 float* __restrict__ delta_out,
                                                                 float* __restrict__ d_enc) {
    using Tl = Tiling<W>;
    constexpr int UB = Tl::UB, QB = Tl::QB, T = Tl::T;
    constexpr int R = W > (int)kChunkK ? W : (int)kChunkK;
    extern __shared__ float4 smem4[];
    float* bufA = reinterpret_cast<float*>(smem4);  // [R][T]
    float* bufB = bufA + R * T;                      // [W][T]
    float* wt = bufB + W * T;                        // [R][W]
    const int tu = threadIdx.x >> 4, tq = threadIdx.x & 15;
    const int u0 = tu * UB, q0 = tq * QB;
    const uint32_t K = a.n_input, L = a.n_hidden_layers;
    const uint32_t n_tiles = (N + T - 1) / T;
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint32_t tq0 = tile * T;
        __syncthreads();
        // output activation transfer (fully_fused_mlp.cu:757-760): dz[r][t] in bufA rows 0..15
        for (uint32_t idx = threadIdx.x; idx < 16u * T; idx += kThreads) {
            const uint32_t r = idx / T, t = idx - r * T, q = tq0 + t;
            float v = 0.0f;
            if (q < N && r < a.n_output)
                v = d_out[(size_t)q * a.n_output + r] * act_bwd(a.out_act, out[(size_t)q * a.n_output + r]);
            bufA[idx] = v;
            if (q < N && r < a.n_output_padded) delta_out[(size_t)r * N + q] = v;
        }
        // W_out [16][W] row-major is already k-major for the transposed product
        stage_rows(wt, w + (size_t)W * K + (size_t)(L - 1) * W * W, a.n_output, W, 0, W, W);
        __syncthreads();
        float acc[UB][QB];
#pragma unroll
        for (int i = 0; i < UB; ++i)
#pragma unroll
            for (int j = 0; j < QB; ++j) acc[i][j] = 0.0f;
        lds_gemm<W>(wt, bufA, a.n_output, u0, q0, acc);
        for (int layer = (int)L - 1; layer >= 0; --layer) {
            // through the hidden activation: delta = acc * act'(h_layer)
#pragma unroll
            for (int i = 0; i < UB; ++i) {
                float hv[QB], dv[QB];
#pragma unroll
                for (int j = 0; j < QB; ++j) {
                    const uint32_t q = tq0 + q0 + j;
                    hv[j] = q < N ? hidden[((size_t)layer * W + u0 + i) * N + q] : 0.0f;
                }
                act_bwd_n<QB>(a.act, hv, dv);
#pragma unroll
                for (int j = 0; j < QB; ++j) {
                    const uint32_t q = tq0 + q0 + j;
                    acc[i][j] *= dv[j];
                    if (q < N) delta[((size_t)layer * W + u0 + i) * N + q] = acc[i][j];
                }
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < UB; ++i)
#pragma unroll
                for (int j = 0; j < QB; ++j) bufB[(u0 + i) * T + q0 + j] = acc[i][j];
            if (layer == 0) break;
            stage_rows(wt, w + (size_t)W * K + (size_t)(layer - 1) * W * W, W, W, 0, W, W);
            __syncthreads();
#pragma unroll
            for (int i = 0; i < UB; ++i)
#pragma unroll
                for (int j = 0; j < QB; ++j) acc[i][j] = 0.0f;
            lds_gemm<W>(wt, bufB, W, u0, q0, acc);
        }
        if (!d_enc) continue;
        // dL/d(enc)[m][t] = sum_u W0[u][m] delta_0[u][t], in chunks of W rows of m
        for (uint32_t m0 = 0; m0 < K; m0 += W) {
            const uint32_t mm = min((uint32_t)W, K - m0);
            __syncthreads();
            stage_rows(wt, w, W, K, m0, mm, W);
            for (uint32_t idx = threadIdx.x; idx < (uint32_t)W * (W - mm); idx += kThreads) {
                const uint32_t k = idx / (W - mm), m = mm + idx % (W - mm);
                wt[k * W + m] = 0.0f;
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < UB; ++i)
#pragma unroll
                for (int j = 0; j < QB; ++j) acc[i][j] = 0.0f;
            lds_gemm<W>(wt, bufB, W, u0, q0, acc);
#pragma unroll
            for (int i = 0; i < UB; ++i)
#pragma unroll
                for (int j = 0; j < QB; ++j) {
                    const uint32_t m = m0 + u0 + i, q = tq0 + q0 + j;
                    if (m < K && q < N) d_enc[(size_t)m * N + q] = acc[i][j];
                }
        }
    }
}

// ---- weight gradients: C[m][k] = sum_n A[m][n] * B[k][n] over a chunk of n ----
constexpr int kTile = 64;
constexpr int kSub = 32;
constexpr uint32_t kNChunk = 1024;

__global__ __launch_bounds__(256) void wgrad_partial_kernel(uint32_t N, uint32_t M, uint32_t Kb,
                                                            const float* __restrict__ A, const float* __restrict__ B,
                                                            float* __restrict__ partial) {
    __shared__ float As[kSub][kTile + 4];
    __shared__ float Bs[kSub][kTile + 4];
    const uint32_t m0 = blockIdx.x * kTile, k0 = blockIdx.y * kTile;
    const uint32_t n_begin = blockIdx.z * kNChunk, n_end = min(N, n_begin + kNChunk);
    const int tm = threadIdx.x / 16, tk = threadIdx.x % 16;
    float acc[4][4] = {};
    for (uint32_t n0 = n_begin; n0 < n_end; n0 += kSub) {
        for (int idx = threadIdx.x; idx < kSub * kTile; idx += 256) {
            const int row = idx / kSub, nn = idx % kSub;
            const uint32_t n = n0 + nn;
            As[nn][row] = (m0 + row < M && n < n_end) ? A[(size_t)(m0 + row) * N + n] : 0.0f;
            Bs[nn][row] = (k0 + row < Kb && n < n_end) ? B[(size_t)(k0 + row) * N + n] : 0.0f;
        }
        __syncthreads();
#pragma unroll 8
        for (int nn = 0; nn < kSub; ++nn) {
            float av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                av[u] = As[nn][tm * 4 + u];
                bv[u] = Bs[nn][tk * 4 + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] += av[u] * bv[v];
        }
        __syncthreads();
    }
    float* p = partial + (size_t)blockIdx.z * M * Kb;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint32_t m = m0 + tm * 4 + u, k = k0 + tk * 4 + v;
            if (m < M && k < Kb) p[(size_t)m * Kb + k] = acc[u][v];
        }
}

__global__ __launch_bounds__(256) void wgrad_reduce_kernel(uint32_t n_elems, uint32_t n_chunks,
                                                           const float* __restrict__ partial, float* __restrict__ out,
                                                           int accumulate) {
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n_elems) return;
    float s = 0.0f;
    for (uint32_t c = 0; c < n_chunks; ++c) s += partial[(size_t)c * n_elems + e];
    out[e] = accumulate ? out[e] + s : s;
}

template <typename K>
void by_width(uint32_t W, K&& k) {
    switch (W) {
        case 16: k(std::integral_constant<int, 16>{}); break;
        case 32: k(std::integral_constant<int, 32>{}); break;
        case 64: k(std::integral_constant<int, 64>{}); break;
        case 128: k(std::integral_constant<int, 128>{}); break;
        default: fail(NLOD_ERR_CONFIG, "unsupported network width");
    }
}

MlpArgs make_args(const NetShape& n) {
    return MlpArgs{n.n_input, n.n_hidden_layers, n.n_output, n.n_output_padded, n.activation, n.output_activation};
}

template <int W>
size_t smem_bytes() {
    constexpr int T = Tiling<W>::T;
    constexpr int R = W > (int)kChunkK ? W : (int)kChunkK;
    return sizeof(float) * ((size_t)R * T + (size_t)W * T + (size_t)R * W);
}

template <int W, typename Kern>
uint32_t grid_for(Kern kern, uint32_t N) {
    const int occ = occupancy_blocks((const void*)kern, kThreads, smem_bytes<W>());
    const uint32_t tiles = div_round_up(N, Tiling<W>::T);
    return std::min<uint32_t>(tiles, (uint32_t)(cu_count() * occ));
}

void wgrad_one(hipStream_t stream, uint32_t N, uint32_t M, uint32_t Kb, const float* A, const float* B, float* out,
               float* ws, bool accumulate) {
    const uint32_t n_chunks = div_round_up(N, kNChunk);
    dim3 grid(div_round_up(M, kTile), div_round_up(Kb, kTile), n_chunks);
    hipLaunchKernelGGL(wgrad_partial_kernel, grid, dim3(256), 0, stream, N, M, Kb, A, B, ws);
    NLOD_LAUNCH_CHECK();
    const uint32_t n = M * Kb;
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(div_round_up(n, 256)), dim3(256), 0, stream, n, n_chunks, ws, out,
                       accumulate ? 1 : 0);
    NLOD_LAUNCH_CHECK();
}

}  // namespace

void mlp_forward_f32(hipStream_t stream, const NetShape& net, const float* weights, const float* enc, uint32_t N,
                     float* out, float* hidden) {
    if (N == 0) return;
    const MlpArgs a = make_args(net);
    by_width(net.width, [&](auto Wc) {
        constexpr int W = decltype(Wc)::value;
        auto kern = mlp_fwd_kernel<W>;
        const uint32_t blocks = grid_for<W>(kern, N);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kThreads), smem_bytes<W>(), stream, N, a, weights, enc, out,
                           hidden);
    });
    NLOD_LAUNCH_CHECK();
}

void mlp_backward_delta_f32(hipStream_t stream, const NetShape& net, const float* weights, const float* hidden,
                            const float* out, const float* d_out, uint32_t N, float* delta, float* delta_out,
                            float* d_enc) {
    if (N == 0) return;
    const MlpArgs a = make_args(net);
    by_width(net.width, [&](auto Wc) {
        constexpr int W = decltype(Wc)::value;
        auto kern = mlp_bwd_delta_kernel<W>;
        const uint32_t blocks = grid_for<W>(kern, N);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kThreads), smem_bytes<W>(), stream, N, a, weights, hidden, out,
                           d_out, delta, delta_out, d_enc);
    });
    NLOD_LAUNCH_CHECK();
}

size_t mlp_weight_grad_workspace_floats(const NetShape& net, uint32_t N) {
    const size_t W = net.width;
    const size_t biggest = std::max({W * net.n_input, W * W, (size_t)net.n_output_padded * W});
    return biggest * div_round_up(std::max<uint32_t>(N, 1), kNChunk);
}

void mlp_weight_grad_f32(hipStream_t stream, const NetShape& net, const float* enc, const float* hidden,
                         const float* delta, const float* delta_out, uint32_t N, float* grads, float* workspace,
                         size_t workspace_floats, bool accumulate) {
    if (workspace_floats < mlp_weight_grad_workspace_floats(net, N)) fail(NLOD_ERR_STATE, "weight-gradient workspace too small");
    const uint32_t W = net.width, L = net.n_hidden_layers;
    if (N == 0) {
        if (!accumulate) fill_f32(stream, grads, net.n_params(), 0.0f);
        return;
    }
    wgrad_one(stream, N, W, net.n_input, delta, enc, grads + net.matrix_offset(0), workspace, accumulate);
    for (uint32_t l = 1; l < L; ++l)
        wgrad_one(stream, N, W, W, delta + (size_t)l * W * N, hidden + (size_t)(l - 1) * W * N,
                  grads + net.matrix_offset(l), workspace, accumulate);
    wgrad_one(stream, N, net.n_output_padded, W, delta_out, hidden + (size_t)(L - 1) * W * N,
              grads + net.matrix_offset(L), workspace, accumulate);
}

}  // namespace nlod
