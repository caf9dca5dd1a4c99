// Plain FP32 form of the fully fused MLP (tiny-cuda-nn src/fully_fused_mlp.cu,
// no biases): forward, backward through the layers, and weight gradients.
// One query per lane; the current layer's weight matrix is staged in LDS and
// read as float4 broadcasts (every lane of a wave reads the same row), the
// activations of a query stay in registers.  Weight gradients are batch
// reductions dW = delta . act^T computed as split-N tiled FP32 GEMMs with a
// deterministic second-pass reduction.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "kernels.h"
#include "nlod_common.h"

namespace nlod {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kChunkK = 64;  // layer-0 columns staged at a time

struct MlpArgs {
    uint32_t n_input, n_hidden_layers, n_output, n_output_padded;
    int32_t act, out_act;
};

template <int W>
__device__ __forceinline__ void stage(float* sw, const float* __restrict__ src, uint32_t rows, uint32_t cols,
                                      uint32_t src_stride) {
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < rows * cols; idx += kThreads) {
        const uint32_t r = idx / cols, c = idx - r * cols;
        sw[idx] = src[(size_t)r * src_stride + c];
    }
    __syncthreads();
}

template <int W>
__global__ __launch_bounds__(kThreads) void mlp_fwd_kernel(uint32_t N, MlpArgs a, const float* __restrict__ w,
                                                           const float* __restrict__ enc, float* __restrict__ out,
                                                           float* __restrict__ hidden) {
    extern __shared__ float4 smem4[];
    float* sw = reinterpret_cast<float*>(smem4);
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    const bool valid = i < N;
    const uint32_t K = a.n_input;

    float acc[W];
#pragma unroll
    for (int r = 0; r < W; ++r) acc[r] = 0.0f;
    for (uint32_t k0 = 0; k0 < K; k0 += kChunkK) {
        const uint32_t kc = min(kChunkK, K - k0);
        stage<W>(sw, w + k0, W, kc, K);
        for (uint32_t c = 0; c < kc; c += 4) {
            float x[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) x[j] = valid ? enc[(size_t)(k0 + c + j) * N + i] : 0.0f;
#pragma unroll
            for (int r = 0; r < W; ++r) {
                const float4 ww = *reinterpret_cast<const float4*>(sw + r * kc + c);
                acc[r] += ww.x * x[0] + ww.y * x[1] + ww.z * x[2] + ww.w * x[3];
            }
        }
    }
    float h[W];
#pragma unroll
    for (int r = 0; r < W; ++r) h[r] = act_fwd(a.act, acc[r]);
    if (hidden && valid) {
#pragma unroll
        for (int r = 0; r < W; ++r) hidden[(size_t)r * N + i] = h[r];
    }
    const float* wl = w + (size_t)W * K;
    for (uint32_t layer = 1; layer < a.n_hidden_layers; ++layer) {
        stage<W>(sw, wl, W, W, W);
#pragma unroll
        for (int r = 0; r < W; ++r) {
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < W; k += 4) {
                const float4 ww = *reinterpret_cast<const float4*>(sw + r * W + k);
                s += ww.x * h[k] + ww.y * h[k + 1] + ww.z * h[k + 2] + ww.w * h[k + 3];
            }
            acc[r] = s;
        }
#pragma unroll
        for (int r = 0; r < W; ++r) h[r] = act_fwd(a.act, acc[r]);
        if (hidden && valid) {
#pragma unroll
            for (int r = 0; r < W; ++r) hidden[((size_t)layer * W + r) * N + i] = h[r];
        }
        wl += (size_t)W * W;
    }
    stage<W>(sw, wl, a.n_output, W, W);
    for (uint32_t r = 0; r < a.n_output; ++r) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < W; k += 4) {
            const float4 ww = *reinterpret_cast<const float4*>(sw + r * W + k);
            s += ww.x * h[k] + ww.y * h[k + 1] + ww.z * h[k + 2] + ww.w * h[k + 3];
        }
        if (valid) out[(size_t)i * a.n_output + r] = act_fwd(a.out_act, s);
    }
}

template <int W>
__global__ __launch_bounds__(kThreads) void mlp_bwd_delta_kernel(uint32_t N, MlpArgs a, const float* __restrict__ w,
                                                                 const float* __restrict__ hidden,
                                                                 const float* __restrict__ out,
                                                                 const float* __restrict__ d_out,
                                                                 float* __restrict__ delta,
                                                                 float* __restrict__ delta_out,
                                                                 float* __restrict__ d_enc) {
    extern __shared__ float4 smem4[];
    float* sw = reinterpret_cast<float*>(smem4);
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    const bool valid = i < N;
    const uint32_t K = a.n_input, L = a.n_hidden_layers;

    // output activation transfer (fully_fused_mlp.cu:757-760)
    float dz[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        float v = 0.0f;
        if (valid && (uint32_t)r < a.n_output)
            v = d_out[(size_t)i * a.n_output + r] * act_bwd(a.out_act, out[(size_t)i * a.n_output + r]);
        dz[r] = v;
        if (valid && (uint32_t)r < a.n_output_padded) delta_out[(size_t)r * N + i] = v;
    }
    const float* w_out = w + (size_t)W * K + (size_t)(L - 1) * W * W;
    stage<W>(sw, w_out, a.n_output, W, W);
    float d[W];
#pragma unroll
    for (int k = 0; k < W; ++k) d[k] = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if ((uint32_t)r < a.n_output) {
#pragma unroll
            for (int k = 0; k < W; k += 4) {
                const float4 ww = *reinterpret_cast<const float4*>(sw + r * W + k);
                d[k] += ww.x * dz[r];
                d[k + 1] += ww.y * dz[r];
                d[k + 2] += ww.z * dz[r];
                d[k + 3] += ww.w * dz[r];
            }
        }
    }
    for (int layer = (int)L - 1; layer >= 0; --layer) {
        // d holds dL/d(post-activation h_layer); transfer through the activation
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const float hv = valid ? hidden[((size_t)layer * W + k) * N + i] : 0.0f;
            d[k] *= act_bwd(a.act, hv);
        }
        if (valid) {
#pragma unroll
            for (int k = 0; k < W; ++k) delta[((size_t)layer * W + k) * N + i] = d[k];
        }
        if (layer == 0) break;
        stage<W>(sw, w + (size_t)W * K + (size_t)(layer - 1) * W * W, W, W, W);
        float dp[W];
#pragma unroll
        for (int k = 0; k < W; ++k) dp[k] = 0.0f;
#pragma unroll
        for (int r = 0; r < W; ++r) {
#pragma unroll
            for (int k = 0; k < W; k += 4) {
                const float4 ww = *reinterpret_cast<const float4*>(sw + r * W + k);
                dp[k] += ww.x * d[r];
                dp[k + 1] += ww.y * d[r];
                dp[k + 2] += ww.z * d[r];
                dp[k + 3] += ww.w * d[r];
            }
        }
#pragma unroll
        for (int k = 0; k < W; ++k) d[k] = dp[k];
    }
    if (!d_enc) return;
    // dL/d(enc) = W0^T delta_0
    for (uint32_t k0 = 0; k0 < K; k0 += kChunkK) {
        const uint32_t kc = min(kChunkK, K - k0);
        stage<W>(sw, w + k0, W, kc, K);
        for (uint32_t c = 0; c < kc; c += 4) {
            float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < W; ++r) {
                const float4 ww = *reinterpret_cast<const float4*>(sw + r * kc + c);
                s[0] += ww.x * d[r];
                s[1] += ww.y * d[r];
                s[2] += ww.z * d[r];
                s[3] += ww.w * d[r];
            }
            if (valid) {
#pragma unroll
                for (int j = 0; j < 4; ++j) d_enc[(size_t)(k0 + c + j) * N + i] = s[j];
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

size_t smem_bytes(const NetShape& n) {
    const size_t W = n.width;
    return sizeof(float) * std::max({W * std::min<size_t>(kChunkK, n.n_input), W * W, (size_t)16 * W});
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
        hipLaunchKernelGGL(mlp_fwd_kernel<W>, dim3(div_round_up(N, kThreads)), dim3(kThreads), smem_bytes(net), stream,
                           N, a, weights, enc, out, hidden);
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
        hipLaunchKernelGGL(mlp_bwd_delta_kernel<W>, dim3(div_round_up(N, kThreads)), dim3(kThreads), smem_bytes(net),
                           stream, N, a, weights, hidden, out, d_out, delta, delta_out, d_enc);
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
    // input layer: delta_0 [W][N] x enc [K][N]
    wgrad_one(stream, N, W, net.n_input, delta, enc, grads + net.matrix_offset(0), workspace, accumulate);
    for (uint32_t l = 1; l < L; ++l)
        wgrad_one(stream, N, W, W, delta + (size_t)l * W * N, hidden + (size_t)(l - 1) * W * N,
                  grads + net.matrix_offset(l), workspace, accumulate);
    wgrad_one(stream, N, net.n_output_padded, W, delta_out, hidden + (size_t)(L - 1) * W * N,
              grads + net.matrix_offset(L), workspace, accumulate);
}

}  // namespace nlod
