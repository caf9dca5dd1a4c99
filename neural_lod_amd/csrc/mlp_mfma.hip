// FP16-MFMA form of the encode + fully fused MLP (gfx950).
//
// Fused inference kernel: one workgroup of 256 lanes (4 waves) handles tiles
// of 256 queries.  Phase 1 — each lane encodes one query (hash-grid gathers
// interpolated in FP32, spherical harmonics, identity) and writes the FP16
// features to an LDS tile [query][K].  Phase 2 — each wave runs its 64 queries
// through the MLP with v_mfma_f32_16x16x32_f16 (FP32 accumulation), weights
// resident in LDS.  Between layers the activations never leave registers: the
// 16x16 accumulator layout (lane = query column n, 4 consecutive unit rows per
// lane group g) is fed straight back as the B operand of the next layer, and
// the next layer's weight columns are pre-permuted (at pack time) to the unit
// order that layout implies.
//
// MFMA operand layout (v_mfma_f32_16x16x32_f16, wave64, g = lane / 16):
//   A: lane holds A[m = lane % 16][k-slot 8g .. 8g+7]
//   B: lane holds B[k-slot 8g .. 8g+7][n = lane % 16]
//   D: lane holds D[m = 4g .. 4g+3][n = lane % 16]
// Within one instruction A and B refer to the same (g, slot) -> k mapping, so
// any consistent permutation of k is exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "device_common.h"
#include "model.h"
#include "nlod_common.h"
#include "sh_basis.h"

namespace nlod {
namespace {

constexpr int kMaxSegs = 8;
constexpr int kBlock = 256;   // lanes per workgroup
constexpr int kTileQ = 256;   // queries per workgroup tile (one per lane in phase 1)

// Encoding description read by the fused kernels (uniform, from device memory).
struct DevSegment {
    int32_t type;
    uint32_t in_offset, out_offset, n_out, n_pad, degree;
    float id_scale, id_offset;
    uint64_t param_offset;
    GridLevels lv;
};
struct EncProgram {
    uint32_t n_segs, n_input_dims, enc_width, k_padded;
    DevSegment segs[kMaxSegs];
};

uint32_t k_padded_of(const NetShape& n) { return (n.n_input + 31) / 32 * 32; }

// Packed fp16 weight fragments: per matrix, per 16-row tile, per 32-column
// chunk, 64 lanes x 8 halves.
size_t frag_count(uint32_t rows, uint32_t cols_padded) { return (size_t)(rows / 16) * (cols_padded / 32); }

size_t packed_halves(const NetShape& n) {
    const uint32_t W = n.width;
    size_t frags = frag_count(W, k_padded_of(n)) + (size_t)(n.n_hidden_layers - 1) * frag_count(W, W) +
                   frag_count(16, W);
    return frags * 64 * 8;
}

// k-slot (g, j) of chunk c of a layer fed by a previous MFMA layer's
// accumulators -> hidden unit index.
__host__ __device__ inline uint32_t permuted_unit(uint32_t c, uint32_t g, uint32_t j) {
    return 32u * c + (j < 4 ? 4u * g + j : 16u + 4u * g + (j - 4u));
}

__global__ void pack_kernel(uint32_t W, uint32_t K, uint32_t Kp, uint32_t L, const float* __restrict__ w,
                            half_t* __restrict__ out, uint32_t total_halves) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total_halves) return;
    const uint32_t j = e & 7u, lane = (e >> 3) & 63u, frag = e >> 9;
    const uint32_t g = lane >> 4, r16 = lane & 15u;
    // locate the matrix of this fragment
    const uint32_t f0 = (W / 16) * (Kp / 32);
    const uint32_t fh = (W / 16) * (W / 32);
    float v = 0.0f;
    if (frag < f0) {
        const uint32_t mt = frag / (Kp / 32), c = frag % (Kp / 32);
        const uint32_t row = 16 * mt + r16, col = 32 * c + 8 * g + j;
        v = col < K ? w[(size_t)row * K + col] : 0.0f;
    } else if (frag < f0 + (L - 1) * fh) {
        const uint32_t f = frag - f0, layer = 1 + f / fh, ff = f % fh;
        const uint32_t mt = ff / (W / 32), c = ff % (W / 32);
        const uint32_t row = 16 * mt + r16, col = permuted_unit(c, g, j);
        v = w[(size_t)W * K + (size_t)(layer - 1) * W * W + (size_t)row * W + col];
    } else {
        const uint32_t c = frag - f0 - (L - 1) * fh;
        const uint32_t row = r16, col = permuted_unit(c, g, j);
        v = w[(size_t)W * K + (size_t)(L - 1) * W * W + (size_t)row * W + col];
    }
    out[e] = (half_t)v;
}

template <int F>
__device__ __forceinline__ void grid_level_f16(const GridLevels& lv, uint32_t level, const float* x,
                                               const float* __restrict__ table, half_t* dst) {
    const float scale = lv.scale[level];
    const uint32_t res = lv.res[level];
    const uint32_t off = lv.offset[level];
    const uint32_t hs = lv.offset[level + 1] - off;
    const float* tab = table + (size_t)off * F;
    float fr[3];
    uint32_t cell[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) grid_pos(x[d], scale, lv.interp, fr[d], cell[d]);
    float acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.0f;
    if (lv.interp == INTERP_NEAREST) {
        const uint32_t idx = grid_index(lv.grid_type, hs, res, cell[0], cell[1], cell[2]);
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = tab[(size_t)idx * F + f];
    } else {
        float vals[8][F];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const uint32_t idx = grid_index(lv.grid_type, hs, res, cell[0] + (c & 1), cell[1] + ((c >> 1) & 1),
                                            cell[2] + ((c >> 2) & 1));
            if constexpr (F == 2) {
                const float2 v2 = *reinterpret_cast<const float2*>(tab + (size_t)idx * 2);
                vals[c][0] = v2.x;
                vals[c][1] = v2.y;
            } else {
#pragma unroll
                for (int f = 0; f < F; ++f) vals[c][f] = tab[(size_t)idx * F + f];
            }
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            float w = 1.0f;
#pragma unroll
            for (int d = 0; d < 3; ++d) w *= ((c >> d) & 1) ? fr[d] : 1.0f - fr[d];
#pragma unroll
            for (int f = 0; f < F; ++f) acc[f] += w * vals[c][f];
        }
    }
#pragma unroll
    for (int f = 0; f < F; ++f) dst[level * F + f] = (half_t)acc[f];
}

// Encodes one query into row `dst` (FP16, natural feature order, zero-padded to k_padded).
__device__ void encode_query_f16(const EncProgram& P, const float* __restrict__ params, const float* q, half_t* dst) {
    for (uint32_t si = 0; si < P.n_segs; ++si) {
        const DevSegment& s = P.segs[si];
        const float* x = q + s.in_offset;
        half_t* o = dst + s.out_offset;
        if (s.type == SEG_GRID) {
            const float* table = params + s.param_offset;
            const uint32_t nl = s.lv.n_levels;
            switch (s.lv.F) {
                case 1: for (uint32_t l = 0; l < nl; ++l) grid_level_f16<1>(s.lv, l, x, table, o); break;
                case 2: for (uint32_t l = 0; l < nl; ++l) grid_level_f16<2>(s.lv, l, x, table, o); break;
                case 4: for (uint32_t l = 0; l < nl; ++l) grid_level_f16<4>(s.lv, l, x, table, o); break;
                default: for (uint32_t l = 0; l < nl; ++l) grid_level_f16<8>(s.lv, l, x, table, o); break;
            }
            for (uint32_t p = 0; p < s.n_pad; ++p) o[s.n_out + p] = (half_t)0.0f;
        } else if (s.type == SEG_SH) {
            const float vx = x[0] * 2.f - 1.f, vy = x[1] * 2.f - 1.f, vz = x[2] * 2.f - 1.f;
            float X[8], Y[8], Z[8];
            X[0] = Y[0] = Z[0] = 1.0f;
#pragma unroll
            for (int k = 1; k < 8; ++k) {
                X[k] = X[k - 1] * vx;
                Y[k] = Y[k - 1] * vy;
                Z[k] = Z[k - 1] * vz;
            }
            for (uint32_t p = 0; p < s.n_pad; ++p) o[p] = (half_t)1.0f;
            half_t* os = o + s.n_pad;
            sh_basis(s.degree, X, Y, Z, [&](int k, float v) { os[k] = (half_t)v; });
        } else {
            for (uint32_t k = 0; k < s.n_out; ++k) o[k] = (half_t)(x[k] * s.id_scale + s.id_offset);
            for (uint32_t p = 0; p < s.n_pad; ++p) o[s.n_out + p] = (half_t)1.0f;
        }
    }
    for (uint32_t k = P.enc_width; k < P.k_padded; ++k) dst[k] = (half_t)0.0f;
}

__device__ __forceinline__ floatx4 mfma16(const half8& a, const half8& b, const floatx4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ half8 pack_act(const floatx4& lo, const floatx4& hi, int act) {
    half8 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        r[i] = (half_t)act_fwd(act, lo[i]);
        r[i + 4] = (half_t)act_fwd(act, hi[i]);
    }
    return r;
}

// Runs the MLP for 16 queries whose FP16 features are rows feat[0..15] of
// stride fstride (halves).  Returns the output-layer accumulator (outputs 4g..4g+3
// for query lane % 16).
template <int W>
__device__ __forceinline__ floatx4 mlp_tile16(const half_t* __restrict__ sw, const half_t* __restrict__ feat,
                                              uint32_t fstride, uint32_t Kp, uint32_t L, int act) {
    constexpr int MT = W / 16, KC = W / 32;
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    floatx4 acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
    const uint32_t kc0 = Kp / 32;
    const half_t* frag = sw;
    for (uint32_t c = 0; c < kc0; ++c) {
        const half8 b = *reinterpret_cast<const half8*>(feat + (size_t)n * fstride + 32 * c + 8 * g);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const half8 a = *reinterpret_cast<const half8*>(frag + ((size_t)(mt * kc0 + c) * 64 + lane) * 8);
            acc[mt] = mfma16(a, b, acc[mt]);
        }
    }
    frag += (size_t)MT * kc0 * 512;
    half8 b[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) b[c] = pack_act(acc[2 * c], acc[2 * c + 1], act);
    for (uint32_t layer = 1; layer < L; ++layer) {
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            floatx4 s = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                const half8 a = *reinterpret_cast<const half8*>(frag + ((size_t)(mt * KC + c) * 64 + lane) * 8);
                s = mfma16(a, b[c], s);
            }
            acc[mt] = s;
        }
        frag += (size_t)MT * KC * 512;
#pragma unroll
        for (int c = 0; c < KC; ++c) b[c] = pack_act(acc[2 * c], acc[2 * c + 1], act);
    }
    floatx4 o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < KC; ++c) {
        const half8 a = *reinterpret_cast<const half8*>(frag + ((size_t)c * 64 + lane) * 8);
        o = mfma16(a, b[c], o);
    }
    return o;
}

struct FusedArgs {
    uint32_t N, n_input_dims, n_output, n_hidden_layers, k_padded, fstride;
    int32_t act, out_act;
    uint32_t packed_halves;
};

template <int W>
__global__ __launch_bounds__(kBlock) void fused_inference_kernel(FusedArgs a, const EncProgram* __restrict__ prog,
                                                                 const float* __restrict__ params,
                                                                 const half_t* __restrict__ packed,
                                                                 const float* __restrict__ in,
                                                                 float* __restrict__ out) {
    extern __shared__ float4 smem4[];
    half_t* sw = reinterpret_cast<half_t*>(smem4);
    half_t* feat = sw + a.packed_halves;
    // weights -> LDS (16-byte copies)
    for (uint32_t k = threadIdx.x; k < a.packed_halves / 8; k += kBlock)
        reinterpret_cast<float4*>(sw)[k] = reinterpret_cast<const float4*>(packed)[k];
    const EncProgram& P = *prog;
    const uint32_t n_tiles = (a.N + kTileQ - 1) / kTileQ;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint32_t q0 = tile * kTileQ;
        __syncthreads();  // previous tile's MLP reads of feat are done (and weights staged)
        {
            const uint32_t q = q0 + threadIdx.x;
            half_t* row = feat + (size_t)threadIdx.x * a.fstride;
            if (q < a.N) encode_query_f16(P, params, in + (size_t)q * a.n_input_dims, row);
            else for (uint32_t k = 0; k < a.k_padded; ++k) row[k] = (half_t)0.0f;
        }
        __syncthreads();
#pragma unroll 1
        for (int t = 0; t < 4; ++t) {
            const uint32_t local = wave * 64 + t * 16;
            if (q0 + local >= a.N) break;
            const floatx4 o = mlp_tile16<W>(sw, feat + (size_t)local * a.fstride, a.fstride, a.k_padded,
                                            a.n_hidden_layers, a.act);
            const uint32_t q = q0 + local + n;
            if (q < a.N) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t oi = 4 * g + i;
                    if (oi < a.n_output) out[(size_t)q * a.n_output + oi] = act_fwd(a.out_act, o[i]);
                }
            }
        }
    }
}

EncProgram build_program(const Layout& L) {
    EncProgram P{};
    P.n_segs = (uint32_t)L.segs.size();
    P.n_input_dims = L.n_input_dims;
    P.enc_width = L.enc_width;
    P.k_padded = k_padded_of(L.net);
    for (size_t i = 0; i < L.segs.size(); ++i) {
        const Segment& s = L.segs[i];
        DevSegment& d = P.segs[i];
        d.type = s.type;
        d.in_offset = s.in_offset;
        d.out_offset = s.out_offset;
        d.n_out = s.n_out;
        d.n_pad = s.n_pad;
        d.degree = s.degree;
        d.id_scale = s.id_scale;
        d.id_offset = s.id_offset;
        d.param_offset = s.param_offset;
        d.lv = s.levels;
    }
    return P;
}

template <typename K>
void by_width(uint32_t W, K&& k) {
    switch (W) {
        case 32: k(std::integral_constant<int, 32>{}); break;
        case 64: k(std::integral_constant<int, 64>{}); break;
        case 128: k(std::integral_constant<int, 128>{}); break;
        default: fail(NLOD_ERR_UNSUPPORTED, "FP16 MFMA kernels support widths 32, 64 and 128");
    }
}

int cu_count() {
    static int n = [] {
        int dev = 0, c = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) return 256;
        return c;
    }();
    return n;
}

}  // namespace

bool mfma_supported(const Layout& L) {
    if (L.net.width != 32 && L.net.width != 64 && L.net.width != 128) return false;
    if (L.segs.size() > (size_t)kMaxSegs) return false;
    if (L.net.n_output > 16) return false;
    if (mfma_fused_smem_bytes(L.net) > 160 * 1024) return false;
    return true;
}

size_t mfma_packed_bytes(const NetShape& net) { return packed_halves(net) * sizeof(half_t); }

size_t mfma_program_bytes() { return sizeof(EncProgram); }

void mfma_upload_program(const Layout& L, void* d_program) {
    const EncProgram host = build_program(L);
    NLOD_HIP(hipMemcpy(d_program, &host, sizeof(EncProgram), hipMemcpyHostToDevice));
}

size_t mfma_fused_smem_bytes(const NetShape& net) {
    const size_t kp = k_padded_of(net);
    return sizeof(half_t) * (packed_halves(net) + (size_t)kTileQ * (kp + 8));
}

void mfma_pack_weights(hipStream_t stream, const NetShape& net, const float* weights, void* packed) {
    const uint32_t total = (uint32_t)packed_halves(net);
    hipLaunchKernelGGL(pack_kernel, dim3(div_round_up(total, 256)), dim3(256), 0, stream, net.width, net.n_input,
                       k_padded_of(net), net.n_hidden_layers, weights, static_cast<half_t*>(packed), total);
    NLOD_LAUNCH_CHECK();
}

void mfma_fused_inference(hipStream_t stream, const Layout& L, const void* d_program, const float* params,
                          const void* packed, const float* in, uint32_t N, float* out) {
    if (N == 0) return;
    const EncProgram* prog = static_cast<const EncProgram*>(d_program);
    FusedArgs a{};
    a.N = N;
    a.n_input_dims = L.n_input_dims;
    a.n_output = L.net.n_output;
    a.n_hidden_layers = L.net.n_hidden_layers;
    a.k_padded = k_padded_of(L.net);
    a.fstride = a.k_padded + 8;  // 16-byte row skew against LDS bank conflicts
    a.act = L.net.activation;
    a.out_act = L.net.output_activation;
    a.packed_halves = (uint32_t)packed_halves(L.net);
    const size_t smem = mfma_fused_smem_bytes(L.net);
    if (smem > 160 * 1024) fail(NLOD_ERR_UNSUPPORTED, "fused kernel LDS footprint exceeds 160 KiB");
    const uint32_t n_tiles = div_round_up(N, kTileQ);
    const uint32_t blocks = std::min<uint32_t>(n_tiles, (uint32_t)cu_count() * 8);
    by_width(L.net.width, [&](auto Wc) {
        constexpr int W = decltype(Wc)::value;
        auto kern = fused_inference_kernel<W>;
        NLOD_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), smem, stream, a, prog, params,
                           static_cast<const half_t*>(packed), in, out);
    });
    NLOD_LAUNCH_CHECK();
}

}  // namespace nlod
