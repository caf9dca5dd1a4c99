// FP16-MFMA training kernels of the fully fused MLP (gfx950): forward with
// saved activations, backward through the layers, and weight gradients.
// Semantics follow tiny-cuda-nn src/fully_fused_mlp.cu:706-833 (no biases).
//
// Saved tensors (per 16-query tile, "unit-major": [tile][unit][16 queries], FP16):
//   H_l   post-activation of hidden layer l (l = 0 .. L-1)
//   D_l   dL/d(pre-activation) of layer l (l = 0 .. L-1), D_L = output layer
// The encoded features X stay in the fragment layout of the encoding pass.
//
// Backward per 16-query tile (one wave): delta_out = dL/dy * act'(y) in the
// accumulator layout, then G_{l-1} = W_l^T delta_l on MFMA with the transposed
// weights pre-packed (columns permuted to the accumulator layout of delta_l),
// delta_{l-1} = G_{l-1} * act'(H_{l-1}); finally dL/d(enc) = W_0^T delta_0.
// Weight gradients dW_l = sum_q delta_l[u][q] H_{l-1}[k][q] use
// v_mfma_f32_16x16x16_f16 with the queries as the reduction dimension, split
// over query chunks with a deterministic second-pass reduction.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "mfma_common.h"
#include "model.h"
#include "nlod_common.h"

namespace nlod {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kWgradTilesPerBlock = 64;  // 1024 queries per weight-gradient partial

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ size_t unit_major(uint32_t tile, uint32_t units, uint32_t u, uint32_t n) {
    return ((size_t)tile * units + u) * 16 + n;
}

__device__ __forceinline__ floatx4 mfma16x16x16(const half4_t& a, const half4_t& b, const floatx4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

struct TrainArgs {
    uint32_t N, n_output, n_hidden_layers, k_enc, k_padded, width;
    int32_t act, out_act;
    uint32_t packed_halves;
};

// Forward for training: like the inference MLP but stores every hidden
// activation (exactly the FP16 values fed to the next layer) and y.
template <int W>
__global__ __launch_bounds__(kBlock) void mlp_fwd_train_kernel(TrainArgs a, const half_t* __restrict__ packed,
                                                               const half_t* __restrict__ feat,
                                                               half_t* __restrict__ H, float* __restrict__ out) {
    constexpr int MT = W / 16, KC = W / 32;
    extern __shared__ float4 smem4[];
    half_t* sw = reinterpret_cast<half_t*>(smem4);
    for (uint32_t k = threadIdx.x; k < a.packed_halves / 8; k += kBlock)
        reinterpret_cast<float4*>(sw)[k] = reinterpret_cast<const float4*>(packed)[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const uint32_t kc0 = a.k_padded / 32, kp2 = a.k_padded / 2;
    const uint32_t n_tiles = (a.N + 15) / 16;
    const uint32_t n_waves = gridDim.x * (kBlock / 64);
    const size_t layer_stride = (size_t)n_tiles * W * 16;
    for (uint32_t tile = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); tile < n_tiles; tile += n_waves) {
        floatx4 acc[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
        for (uint32_t c = 0; c < kc0; ++c) {
            const half8 b = load_feat_frag(feat, tile, kp2, c, g, n);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const half8 w = *reinterpret_cast<const half8*>(sw + ((size_t)(mt * kc0 + c) * 64 + lane) * 8);
                acc[mt] = mfma16(w, b, acc[mt]);
            }
        }
        const half_t* frag = sw + (size_t)MT * kc0 * 512;
        half8 b[KC];
        for (uint32_t layer = 0;; ++layer) {
            half_t* Hl = H + layer * layer_stride;
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                b[c] = pack_act(acc[2 * c], acc[2 * c + 1], a.act);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    Hl[unit_major(tile, W, 32 * c + 4 * g + i, n)] = b[c][i];
                    Hl[unit_major(tile, W, 32 * c + 16 + 4 * g + i, n)] = b[c][i + 4];
                }
            }
            if (layer + 1 >= a.n_hidden_layers) break;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                floatx4 s = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int c = 0; c < KC; ++c) {
                    const half8 w = *reinterpret_cast<const half8*>(frag + ((size_t)(mt * KC + c) * 64 + lane) * 8);
                    s = mfma16(w, b[c], s);
                }
                acc[mt] = s;
            }
            frag += (size_t)MT * KC * 512;
        }
        floatx4 o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            const half8 w = *reinterpret_cast<const half8*>(frag + ((size_t)c * 64 + lane) * 8);
            o = mfma16(w, b[c], o);
        }
        const uint32_t q = tile * 16 + n;
        if (q < a.N) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t oi = 4 * g + i;
                if (oi < a.n_output) out[(size_t)q * a.n_output + oi] = act_fwd(a.out_act, o[i]);
            }
        }
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void mlp_bwd_kernel(TrainArgs a, const half_t* __restrict__ packed_bwd,
                                                         const half_t* __restrict__ H, const float* __restrict__ y,
                                                         const float* __restrict__ dy, half_t* __restrict__ D,
                                                         float* __restrict__ d_enc) {
    constexpr int MT = W / 16, KC = W / 32;
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const uint32_t n_tiles = (a.N + 15) / 16;
    const uint32_t n_waves = gridDim.x * (kBlock / 64);
    const uint32_t L = a.n_hidden_layers;
    const size_t layer_stride = (size_t)n_tiles * W * 16;
    half_t* D_out = D + (size_t)L * layer_stride;
    for (uint32_t tile = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); tile < n_tiles; tile += n_waves) {
        const uint32_t q = tile * 16 + n;
        const bool valid = q < a.N;
        // delta_out = dL/dy * act'(y) (fully_fused_mlp.cu:757-760), rows 4g+i = outputs
        half8 bo;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t oi = 4 * g + i;
            float d = 0.0f;
            if (valid && oi < a.n_output)
                d = dy[(size_t)q * a.n_output + oi] * act_bwd(a.out_act, y[(size_t)q * a.n_output + oi]);
            bo[i] = (half_t)d;
            bo[i + 4] = (half_t)0.0f;
            D_out[unit_major(tile, 16, oi, n)] = bo[i];
        }
        floatx4 acc[MT];
        const half_t* frag = packed_bwd;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const half8 w = *reinterpret_cast<const half8*>(frag + ((size_t)mt * 64 + lane) * 8);
            acc[mt] = mfma16(w, bo, floatx4{0.f, 0.f, 0.f, 0.f});
        }
        frag += (size_t)MT * 512;
        half8 b[KC];
        for (int layer = (int)L - 1; layer >= 0; --layer) {
            const half_t* Hl = H + (size_t)layer * layer_stride;
            half_t* Dl = D + (size_t)layer * layer_stride;
            // delta_layer = G * act'(H_layer), stored unit-major, packed as next B fragments
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                float y[8], dv[8];  // units 32c + 4g + i (i < 4) and 16 above them
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t u0 = 32 * c + 4 * g + i;
                    y[i] = (float)Hl[unit_major(tile, W, u0, n)];
                    y[i + 4] = (float)Hl[unit_major(tile, W, u0 + 16, n)];
                }
                act_bwd_n<8>(a.act, y, dv);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t u0 = 32 * c + 4 * g + i, u1 = u0 + 16;
                    const half_t d0 = (half_t)(acc[2 * c][i] * dv[i]);
                    const half_t d1 = (half_t)(acc[2 * c + 1][i] * dv[i + 4]);
                    b[c][i] = d0;
                    b[c][i + 4] = d1;
                    Dl[unit_major(tile, W, u0, n)] = d0;
                    Dl[unit_major(tile, W, u1, n)] = d1;
                }
            }
            if (layer == 0) break;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                floatx4 s = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int c = 0; c < KC; ++c) {
                    const half8 w = *reinterpret_cast<const half8*>(frag + ((size_t)(mt * KC + c) * 64 + lane) * 8);
                    s = mfma16(w, b[c], s);
                }
                acc[mt] = s;
            }
            frag += (size_t)MT * KC * 512;
        }
        if (!d_enc) continue;
        // dL/d(enc) = W_0^T delta_0, rows = encoded features
        const uint32_t mt0 = a.k_padded / 16;
        for (uint32_t mt = 0; mt < mt0; ++mt) {
            floatx4 s = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                const half8 w = *reinterpret_cast<const half8*>(frag + ((size_t)(mt * KC + c) * 64 + lane) * 8);
                s = mfma16(w, b[c], s);
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
}

// dW partials: grid (chunks, L+1).  Matrix l: rows = units of layer l's output
// (16 for the output layer), cols = inputs (K for layer 0).  Wave w of a block
// owns output tiles t = w, w+4, ... of the matrix and reduces over the chunk's
// query tiles.
template <int W>
__global__ __launch_bounds__(kBlock) void wgrad_mfma_kernel(TrainArgs a, const half_t* __restrict__ feat,
                                                            const half_t* __restrict__ H,
                                                            const half_t* __restrict__ D,
                                                            float* __restrict__ partial, uint32_t partial_stride) {
    constexpr int MAXT = (W / 16) * (W / 16) / 4 > 0 ? (W / 16) * (W / 16) / 4 : 1;
    constexpr int TPW = MAXT < 1 ? 1 : MAXT;  // tiles per wave upper bound for hidden layers
    const int lane = threadIdx.x & 63, g = lane >> 4, m16 = lane & 15, wave = threadIdx.x >> 6;
    const uint32_t L = a.n_hidden_layers;
    const uint32_t layer = blockIdx.y;  // 0 .. L
    const uint32_t n_tiles = (a.N + 15) / 16;
    const size_t layer_stride = (size_t)n_tiles * W * 16;
    const uint32_t rows = layer == L ? 16u : (uint32_t)W;
    const uint32_t cols = layer == 0 ? a.k_padded : (uint32_t)W;
    const uint32_t rt = rows / 16, ct = cols / 16, n_out_tiles = rt * ct;
    const uint32_t kp2 = a.k_padded / 2;
    const half_t* Dl = D + (size_t)layer * layer_stride;  // layer L: the output deltas
    const uint32_t d_units = layer == L ? 16u : (uint32_t)W;
    const half_t* Hp = layer == 0 ? nullptr : H + (size_t)(layer - 1) * layer_stride;
    // up to 4 * TPW output tiles per block pass; loop if the matrix has more (layer 0 with a wide K)
    for (uint32_t t0 = 0; t0 < n_out_tiles; t0 += 4 * TPW) {
        floatx4 acc[TPW];
#pragma unroll
        for (int t = 0; t < TPW; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
        const uint32_t tile_begin = blockIdx.x * kWgradTilesPerBlock;
        const uint32_t tile_end = min(n_tiles, tile_begin + kWgradTilesPerBlock);
        for (uint32_t qt = tile_begin; qt < tile_end; ++qt) {
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
                const uint32_t ot = t0 + (uint32_t)wave + 4u * (uint32_t)t;
                if (ot >= n_out_tiles) break;
                const uint32_t mr = ot / ct, nc = ot % ct;
                // A = delta[u = 16 mr + m16][q = 4g..4g+3]
                const half4_t av = *reinterpret_cast<const half4_t*>(Dl + unit_major(qt, d_units, 16 * mr + m16, 4 * g));
                half4_t bv;
                const uint32_t k = 16 * nc + m16;
                if (layer == 0) {  // B = X[k][q], X in the pair layout
                    const half_t* fb = feat + feat_pair_offset(qt, k >> 1, kp2, 0) + (k & 1);
#pragma unroll
                    for (int j = 0; j < 4; ++j) bv[j] = fb[(4 * g + j) * 2];
                } else {
                    bv = *reinterpret_cast<const half4_t*>(Hp + unit_major(qt, W, k, 4 * g));
                }
                acc[t] = mfma16x16x16(av, bv, acc[t]);
            }
        }
        // write this chunk's partial (row-major [rows][cols] of the matrix)
        float* p = partial + (size_t)blockIdx.x * partial_stride;
        const size_t moff = layer == 0 ? 0 : (size_t)W * a.k_enc + (size_t)(layer - 1) * W * W;
        const uint32_t ld = layer == 0 ? a.k_enc : (uint32_t)W;
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
            const uint32_t ot = t0 + (uint32_t)wave + 4u * (uint32_t)t;
            if (ot >= n_out_tiles) break;
            const uint32_t mr = ot / ct, nc = ot % ct;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t r = 16 * mr + 4 * g + i, c = 16 * nc + m16;
                if (c < ld) p[moff + (size_t)r * ld + c] = acc[t][i];
            }
        }
    }
}

__global__ __launch_bounds__(256) void wgrad_reduce_kernel(uint32_t n_elems, uint32_t n_chunks, uint32_t stride,
                                                           const float* __restrict__ partial, float* __restrict__ out,
                                                           int accumulate) {
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n_elems) return;
    float s = 0.0f;
    for (uint32_t c = 0; c < n_chunks; ++c) s += partial[(size_t)c * stride + e];
    out[e] = accumulate ? out[e] + s : s;
}

TrainArgs make_targs(const Layout& L, uint32_t N) {
    TrainArgs a{};
    a.N = N;
    a.n_output = L.net.n_output;
    a.n_hidden_layers = L.net.n_hidden_layers;
    a.k_enc = L.net.n_input;
    a.k_padded = k_padded_of(L.net);
    a.width = L.net.width;
    a.act = L.net.activation;
    a.out_act = L.net.output_activation;
    a.packed_halves = (uint32_t)packed_halves(L.net);
    return a;
}

template <typename Fn>
void by_w(uint32_t W, Fn&& f) {
    switch (W) {
        case 32: f(std::integral_constant<int, 32>{}); break;
        case 64: f(std::integral_constant<int, 64>{}); break;
        case 128: f(std::integral_constant<int, 128>{}); break;
        default: fail(NLOD_ERR_UNSUPPORTED, "FP16 MFMA kernels support widths 32, 64 and 128");
    }
}

template <typename Kern>
uint32_t persistent_blocks(Kern kern, size_t smem, uint32_t n_tiles) {
    const int occ = occupancy_blocks((const void*)kern, kBlock, smem);
    return std::min<uint32_t>(div_round_up(n_tiles, kBlock / 64), (uint32_t)(cu_count() * occ));
}

}  // namespace

size_t mfma_packed_bwd_bytes(const NetShape& net) { return packed_bwd_halves(net) * sizeof(half_t); }


size_t mfma_hidden_bytes(const Layout& L, uint32_t N) {
    return sizeof(half_t) * (size_t)L.net.n_hidden_layers * div_round_up(N, 16) * L.net.width * 16;
}

size_t mfma_delta_bytes(const Layout& L, uint32_t N) {
    return mfma_hidden_bytes(L, N) + sizeof(half_t) * (size_t)div_round_up(N, 16) * 16 * 16;
}

void mfma_mlp_forward_train(hipStream_t stream, const Layout& L, const void* packed, const void* feat, uint32_t N,
                            void* hidden, float* out) {
    if (N == 0) return;
    const TrainArgs a = make_targs(L, N);
    const size_t smem = mfma_packed_bytes(L.net);
    by_w(L.net.width, [&](auto Wc) {
        constexpr int W = decltype(Wc)::value;
        auto kern = mlp_fwd_train_kernel<W>;
        const uint32_t blocks = persistent_blocks(kern, smem, div_round_up(N, 16));
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), smem, stream, a, static_cast<const half_t*>(packed),
                           static_cast<const half_t*>(feat), static_cast<half_t*>(hidden), out);
    });
    NLOD_LAUNCH_CHECK();
}

void mfma_mlp_backward(hipStream_t stream, const Layout& L, const void* packed_bwd, const void* hidden,
                       const float* y, const float* dy, uint32_t N, void* delta, float* d_enc) {
    if (N == 0) return;
    const TrainArgs a = make_targs(L, N);
    by_w(L.net.width, [&](auto Wc) {
        constexpr int W = decltype(Wc)::value;
        auto kern = mlp_bwd_kernel<W>;
        const uint32_t blocks = persistent_blocks(kern, 0, div_round_up(N, 16));
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), 0, stream, a, static_cast<const half_t*>(packed_bwd),
                           static_cast<const half_t*>(hidden), y, dy, static_cast<half_t*>(delta), d_enc);
    });
    NLOD_LAUNCH_CHECK();
}

size_t mfma_wgrad_workspace_floats(const Layout& L, uint32_t N) {
    const uint32_t chunks = div_round_up(div_round_up(std::max<uint32_t>(N, 1), 16), kWgradTilesPerBlock);
    return (size_t)chunks * L.net.n_params();
}

void mfma_weight_grad(hipStream_t stream, const Layout& L, const void* feat, const void* hidden, const void* delta,
                      uint32_t N, float* grads, float* workspace, size_t workspace_floats, bool accumulate) {
    if (workspace_floats < mfma_wgrad_workspace_floats(L, N)) fail(NLOD_ERR_STATE, "weight-gradient workspace too small");
    const uint32_t n_net = (uint32_t)L.net.n_params();
    if (N == 0) {
        if (!accumulate) fill_f32(stream, grads, n_net, 0.0f);
        return;
    }
    const TrainArgs a = make_targs(L, N);
    const uint32_t chunks = div_round_up(div_round_up(N, 16), kWgradTilesPerBlock);
    by_w(L.net.width, [&](auto Wc) {
        constexpr int W = decltype(Wc)::value;
        hipLaunchKernelGGL(wgrad_mfma_kernel<W>, dim3(chunks, L.net.n_hidden_layers + 1), dim3(kBlock), 0, stream, a,
                           static_cast<const half_t*>(feat), static_cast<const half_t*>(hidden),
                           static_cast<const half_t*>(delta), workspace, n_net);
    });
    NLOD_LAUNCH_CHECK();
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(div_round_up(n_net, 256)), dim3(256), 0, stream, n_net, chunks, n_net,
                       workspace, grads, accumulate ? 1 : 0);
    NLOD_LAUNCH_CHECK();
}

}  // namespace nlod
