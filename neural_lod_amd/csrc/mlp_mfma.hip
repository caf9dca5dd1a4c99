// FP16-MFMA form of the encode + fully fused MLP (gfx950).
//
// Fused inference kernel: every wave independently takes tiles of 16 queries.
// The encoded features never touch memory: for K-chunk c of the first layer,
// lane (n = lane % 16, g = lane / 16) computes exactly the 8 features
// 32c + 8g .. 32c + 8g + 7 of query n that v_mfma_f32_16x16x32_f16 wants in
// its B operand (hash-grid levels gathered and interpolated in FP32, then
// rounded to FP16).  Hidden activations also stay in registers: the 16x16
// accumulator layout (4 consecutive unit rows per lane group) is fed straight
// back as the next layer's B operand, with that layer's weight columns
// pre-permuted at pack time.  Weights (FP16 A fragments) live in LDS, one copy
// per workgroup.
//
// MFMA operand layout (v_mfma_f32_16x16x32_f16, wave64, g = lane / 16):
//   A: lane holds A[m = lane % 16][k-slot 8g .. 8g+7]
//   B: lane holds B[k-slot 8g .. 8g+7][n = lane % 16]
//   D: lane holds D[m = 4g .. 4g+3][n = lane % 16]
// Within one instruction A and B refer to the same (g, slot) -> k mapping, so
// any consistent permutation of k is exact.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "mfma_common.h"
#include "model.h"
#include "nlod_common.h"

namespace nlod {
namespace {

constexpr int kBlock = 256;  // lanes per workgroup (4 waves)

struct FusedArgs {
    uint32_t N, n_input_dims, n_output, n_hidden_layers, k_padded;
    int32_t act, out_act;
    uint32_t packed_halves;
};

template <int W, int F, int MAXD>
__global__ __launch_bounds__(kBlock) void fused_inference_kernel(FusedArgs a, const EncProgram* __restrict__ prog,
                                                                 const float* __restrict__ params,
                                                                 const half_t* __restrict__ packed,
                                                                 const float* __restrict__ in,
                                                                 float* __restrict__ out) {
    constexpr int MT = W / 16;
    extern __shared__ float4 smem4[];
    half_t* sw = reinterpret_cast<half_t*>(smem4);
    for (uint32_t k = threadIdx.x; k < a.packed_halves / 8; k += kBlock)
        reinterpret_cast<float4*>(sw)[k] = reinterpret_cast<const float4*>(packed)[k];
    __syncthreads();
    const EncProgram& P = *prog;
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const uint32_t kc0 = a.k_padded / 32;
    const uint32_t n_tiles = (a.N + 15) / 16;
    const uint32_t n_waves = gridDim.x * (kBlock / 64);
    for (uint32_t tile = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); tile < n_tiles; tile += n_waves) {
        const uint32_t q = tile * 16 + n;
        const float* qin = in + (size_t)min(q, a.N - 1) * a.n_input_dims;
        floatx4 acc[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
        for (uint32_t c = 0; c < kc0; ++c) {
            const half8 b = encode_window<F, MAXD>(P, params, qin, 32 * c + 8 * g);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const half8 w = *reinterpret_cast<const half8*>(sw + ((size_t)(mt * kc0 + c) * 64 + lane) * 8);
                acc[mt] = mfma16(w, b, acc[mt]);
            }
        }
        const floatx4 o = mlp_tail16<W>(sw + (size_t)MT * kc0 * 512, acc, a.n_hidden_layers, a.act);
        if (q < a.N) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t oi = 4 * g + i;
                if (oi < a.n_output) out[(size_t)q * a.n_output + oi] = act_fwd(a.out_act, o[i]);
            }
        }
    }
}

template <int W, int F, int MAXD>
void launch_fused(hipStream_t stream, const FusedArgs& a, const EncProgram* prog, const float* params,
                  const half_t* packed, const float* in, float* out, size_t smem) {
    auto kern = fused_inference_kernel<W, F, MAXD>;
    static int occupancy = -1;
    if (occupancy < 0) {
        NLOD_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        int blocks_per_cu = 0;
        NLOD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, kern, kBlock, smem));
        occupancy = std::max(1, blocks_per_cu);
    }
    const uint32_t n_tiles = div_round_up(a.N, 16);
    const uint32_t blocks = std::min<uint32_t>(div_round_up(n_tiles, kBlock / 64), (uint32_t)(cu_count() * occupancy));
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), smem, stream, a, prog, params, packed, in, out);
    NLOD_LAUNCH_CHECK();
}

}  // namespace

size_t mfma_fused_smem_bytes(const NetShape& net) { return mfma_packed_bytes(net); }

bool mfma_supported(const Layout& L) {
    if (L.net.width != 32 && L.net.width != 64 && L.net.width != 128) return false;
    if (L.segs.size() > (size_t)kMaxSegs) return false;
    if (L.net.n_output > 16) return false;
    if (mfma_fused_smem_bytes(L.net) > 160 * 1024) return false;
    uint32_t F = 0;
    for (const Segment& s : L.segs) {
        if (s.type != SEG_GRID) continue;
        if (F && s.F != F) return false;  // one feature count per model in the fused kernels
        F = s.F;
    }
    return true;
}

void mfma_fused_inference(hipStream_t stream, const Layout& L, const void* d_program, const float* params,
                          const void* packed, const float* in, uint32_t N, float* out) {
    if (N == 0) return;
    FusedArgs a{};
    a.N = N;
    a.n_input_dims = L.n_input_dims;
    a.n_output = L.net.n_output;
    a.n_hidden_layers = L.net.n_hidden_layers;
    a.k_padded = k_padded_of(L.net);
    a.act = L.net.activation;
    a.out_act = L.net.output_activation;
    a.packed_halves = (uint32_t)packed_halves(L.net);
    const size_t smem = mfma_fused_smem_bytes(L.net);
    const EncProgram* prog = static_cast<const EncProgram*>(d_program);
    const half_t* pk = static_cast<const half_t*>(packed);
    dispatch_fused(L, [&](auto Wc, auto Fc, auto Dc) {
        launch_fused<decltype(Wc)::value, decltype(Fc)::value, decltype(Dc)::value>(stream, a, prog, params, pk, in,
                                                                                    out, smem);
    });
}

}  // namespace nlod
