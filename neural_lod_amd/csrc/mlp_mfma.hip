// FP16-MFMA form of the encode + fully fused MLP (gfx950).
//
// Three launches per batch:
//  0. The coarse levels whose FP16 tables fit in LDS together (72 KB: levels 0
//     and 1 of the 16^3-based grids): encode_lds_kernel gathers from a copy of
//     those tables in LDS, sparing the L2 that the other levels saturate.
//     It also copies the grid positions to a [3][N] array (when every grid
//     segment reads the same 3 inputs): each level row then reads 3 lines per
//     32 queries instead of the 6 that 32 rows of the input span (-1.4% for
//     the pass; each level row re-reads the positions).
//  1. The other levels.  Two lanes per (query, grid level), one for each x-side
//     of the cell (grid_level_h16_pair: the two corners of an x-edge then cost
//     one L2 request), blockIdx.y = level, so the workgroups resident at any
//     time all gather from one level's table: that level (2 MiB as FP16 at 2^19
//     entries x 2 features) stays in each XCD's 4 MiB L2 while 2^20 queries
//     stream through it (the reference's "one level per block" idea,
//     grid.h:1053-1058).  The last blockIdx.y row writes the non-grid features
//     (spherical harmonics, identity, padding), overlapping the level rows.
//     Features are stored FP16 as [N/16][K/2][16 queries][2 features]: one
//     level of a 16-query tile is one contiguous 64-byte row, written whole
//     (a layout with 8 features per query row was written a quarter line at a
//     time by four different levels: 3.3x the useful bytes reached HBM).
//     Phase 2 reads a B fragment as 4 dwords, each coalesced across 16 lanes.
//  2. MLP.  Every wave takes tiles of 16 queries; all layers run on
//     v_mfma_f32_16x16x32_f16 with FP32 accumulation, weights resident in
//     LDS, activations never leaving registers: the 16x16 accumulator layout
//     (4 consecutive unit rows per lane group) is fed straight back as the next
//     layer's B operand, with that layer's weight columns pre-permuted at pack
//     time.
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

constexpr int kBlock = 256;
constexpr int kMaxLevelRows = 64;  // grid levels over all grid segments

struct LevelMap {
    uint32_t n;                    // grid levels (blockIdx.y rows before the non-grid row)
    uint8_t seg[kMaxLevelRows];
    uint8_t level[kMaxLevelRows];
};

// Grid levels whose FP16 tables together fit in LDS (the coarse dense levels):
// encoded by encode_lds_kernel instead of rows of encode_kernel.
constexpr int kMaxLdsLevels = 4;
constexpr uint32_t kLdsTableBytes = 72 * 1024;  // two blocks per CU
constexpr int kLdsBlock = 1024;
struct LdsLevels {
    uint32_t n;
    uint32_t halves;                  // FP16 values staged (all levels)
    uint8_t seg[kMaxLdsLevels];
    uint8_t level[kMaxLdsLevels];
    uint32_t at[kMaxLdsLevels];       // first staged half of each level
};

// pair of 2 halves holding feature f of query q
__device__ __forceinline__ uint32_t* pair_ptr(half_t* feat, uint32_t q, uint32_t f, uint32_t kp2) {
    return reinterpret_cast<uint32_t*>(feat + feat_pair_offset(q >> 4, f >> 1, kp2, q & 15));
}

__device__ __forceinline__ uint32_t pack2(half_t a, half_t b) {
    return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}

// Non-grid rows of query q, each computed once and written in the pair layout:
// SH (pad rows of 1.0 first, then the basis), identity (inputs scaled, then pad
// rows of 1.0), the pad rows after a grid (0) and the K tail (0).  Same values,
// bit for bit, as encode_window_f32 gives these rows.
template <int MAXD>
__device__ __forceinline__ void encode_nongrid(const EncProgram& P, const float* __restrict__ x, uint32_t q,
                                               bool valid, half_t* __restrict__ feat) {
#pragma clang fp contract(off)
    // row r of query q sits at base + (r / 2) * 32 + r % 2 (feat_pair_offset)
    half_t* base = feat + feat_pair_offset(q >> 4, 0, P.k_padded / 2, q & 15);
    // Each value is rounded to FP32 first, then to FP16, as everywhere else: the
    // empty asm keeps the compiler from folding the product that made it into
    // the conversion (v_fma_mix, a single rounding).  It emits no instruction.
    auto h = [&](float v) -> half_t {
        asm volatile("" : "+v"(v));
        return valid ? (half_t)v : (half_t)0.0f;
    };
    auto put = [&](uint32_t r, float v) { base[(r >> 1) * 32 + (r & 1)] = h(v); };
    auto put2 = [&](uint32_t r, float a, float b) {  // r even
        *reinterpret_cast<uint32_t*>(base + (r >> 1) * 32) = pack2(h(a), h(b));
    };
    for (uint32_t si = 0; si < P.n_segs; ++si) {
        const DevSegment& s = P.segs[si];
        const uint32_t r0 = s.out_offset;
        if (s.type == SEG_GRID) {
            for (uint32_t k = 0; k < s.n_pad; ++k) put(r0 + s.n_out + k, 0.0f);
            continue;
        }
        const float* xs = x + s.in_offset;
        if (s.type == SEG_SH) {
            for (uint32_t k = 0; k < s.n_pad; ++k) put(r0 + k, 1.0f);
            const uint32_t sh0 = r0 + s.n_pad, nsh = s.degree * s.degree;
            const float vx = xs[0] * 2.f - 1.f, vy = xs[1] * 2.f - 1.f, vz = xs[2] * 2.f - 1.f;
            float X[MAXD], Y[MAXD], Z[MAXD];
            X[0] = Y[0] = Z[0] = 1.0f;
#pragma unroll
            for (int k = 1; k < MAXD; ++k) {
                X[k] = X[k - 1] * vx;
                Y[k] = Y[k - 1] * vy;
                Z[k] = Z[k - 1] * vz;
            }
            float v[MAXD * MAXD];
#pragma unroll
            for (int k = 0; k < MAXD * MAXD; ++k) v[k] = 0.0f;
            sh_basis<MAXD>(s.degree, X, Y, Z, [&](int k, float val) { v[k] = val; });
            if ((sh0 & 1u) == 0) {
#pragma unroll
                for (int k = 0; k < MAXD * MAXD; k += 2) {
                    if ((uint32_t)k + 1 < nsh) put2(sh0 + k, v[k], v[k + 1]);
                    else if ((uint32_t)k < nsh) put(sh0 + k, v[k]);
                }
            } else {
#pragma unroll
                for (int k = 0; k < MAXD * MAXD; ++k)
                    if ((uint32_t)k < nsh) put(sh0 + k, v[k]);
            }
        } else {  // identity
            for (uint32_t k = 0; k < s.n_out; ++k) put(r0 + k, xs[k] * s.id_scale + s.id_offset);
            for (uint32_t k = 0; k < s.n_pad; ++k) put(r0 + s.n_out + k, 1.0f);
        }
    }
    uint32_t r = P.enc_width;
    if (r & 1u) put(r++, 0.0f);
    for (; r + 1 < P.k_padded; r += 2) put2(r, 0.0f, 0.0f);
    if (r < P.k_padded) put(r, 0.0f);
}

template <int F, int MAXD, typename TT>
__global__ __launch_bounds__(kBlock) void encode_kernel(uint32_t N, uint32_t N16, LevelMap map,
                                                        const EncProgram* __restrict__ prog,
                                                        const TT* __restrict__ table, const float* __restrict__ in,
                                                        half_t* __restrict__ feat, const float* __restrict__ xsoa) {
    static_assert(std::is_same<TT, half_t>::value, "the encoding pass gathers from the FP16 table copy");
    const EncProgram& P = *prog;
    const uint32_t kp2 = P.k_padded / 2;
    if (blockIdx.y < map.n) {
        // grid rows: two lanes per query (grid_level_h16_pair); N16 is a multiple
        // of 16, so both lanes of a pair return together
        const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
        const uint32_t q = t >> 1, lane_s = t & 1u;
        if (q >= N16) return;
        const bool valid = q < N;
        const float* x = in + (size_t)(valid ? q : 0) * P.n_input_dims;
        const DevSegment& s = P.segs[map.seg[blockIdx.y]];
        const uint32_t level = map.level[blockIdx.y];
        constexpr int VF = F == 1 ? 1 : F / 2;
        float v[VF];
        float x0, x1, x2;
        if (xsoa) {
            x0 = xsoa[q];
            x1 = xsoa[N16 + q];
            x2 = xsoa[2 * (size_t)N16 + q];
        } else {
            const float* xs = x + s.in_offset;
            x0 = xs[0];
            x1 = xs[1];
            x2 = xs[2];
        }
        grid_level_h16_pair<F>(s.lv, level, x0, x1, x2, table + s.param_offset, lane_s, v);
        const uint32_t f0 = s.out_offset + level * F;  // multiple of F
        if constexpr (F == 1) {
            if (lane_s == 0)
                feat[feat_pair_offset(q >> 4, f0 >> 1, kp2, q & 15) + (f0 & 1)] = valid ? (half_t)v[0] : (half_t)0.0f;
        } else {
            // lane s writes element s of each of the level's feature pairs
#pragma unroll
            for (int k = 0; k < VF; ++k)
                feat[feat_pair_offset(q >> 4, (f0 >> 1) + k, kp2, q & 15) + lane_s] =
                    valid ? (half_t)v[k] : (half_t)0.0f;
        }
        return;
    }
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= N16) return;
    const bool valid = q < N;
    encode_nongrid<MAXD>(P, in + (size_t)(valid ? q : 0) * P.n_input_dims, q, valid, feat);
}

// The levels of `lv` (at most kMaxLdsLevels, tables in LDS): one block copies
// their FP16 tables into LDS once, then each lane encodes a grid-stride run of
// queries, reading the query's position once for all of them.  The gathers come
// from LDS instead of the L2, which every other level row keeps busy.
template <int F>
__global__ __launch_bounds__(kLdsBlock) void encode_lds_kernel(uint32_t N, uint32_t N16, LdsLevels lv,
                                                               const EncProgram* __restrict__ prog,
                                                               const half_t* __restrict__ table,
                                                               const float* __restrict__ in,
                                                               half_t* __restrict__ feat, float* __restrict__ xsoa,
                                                               uint32_t soa_off) {
    extern __shared__ uint4 lds_tab4[];
    half_t* lt = reinterpret_cast<half_t*>(lds_tab4);
    const EncProgram& P = *prog;
    for (uint32_t l = 0; l < lv.n; ++l) {
        const DevSegment& s = P.segs[lv.seg[l]];
        const GridLevels& g = s.lv;
        const uint32_t n = (g.offset[lv.level[l] + 1] - g.offset[lv.level[l]]) * F;
        const half_t* src = table + s.param_offset + (size_t)g.offset[lv.level[l]] * F;
        if ((reinterpret_cast<uintptr_t>(src) & 15u) == 0 && lv.at[l] % 8 == 0 && n % 8 == 0) {
            const uint4* s4 = reinterpret_cast<const uint4*>(src);
            uint4* d4 = reinterpret_cast<uint4*>(lt + lv.at[l]);
            for (uint32_t k = threadIdx.x; k < n / 8; k += kLdsBlock) d4[k] = s4[k];
        } else {
            for (uint32_t k = threadIdx.x; k < n; k += kLdsBlock) lt[lv.at[l] + k] = src[k];
        }
    }
    __syncthreads();
    const uint32_t kp2 = P.k_padded / 2;
    for (uint32_t q = blockIdx.x * kLdsBlock + threadIdx.x; q < N16; q += gridDim.x * kLdsBlock) {
        const bool valid = q < N;
        const float* x = in + (size_t)(valid ? q : 0) * P.n_input_dims;
        if (xsoa) {
            xsoa[q] = x[soa_off];
            xsoa[N16 + q] = x[soa_off + 1];
            xsoa[2 * (size_t)N16 + q] = x[soa_off + 2];
        }
        for (uint32_t l = 0; l < lv.n; ++l) {
            const DevSegment& s = P.segs[lv.seg[l]];
            const uint32_t level = lv.level[l];
            const float* xs = x + s.in_offset;
            float v[F];
            grid_level_at<F, half_t>(s.lv, level, xs[0], xs[1], xs[2], lt + lv.at[l], v);
            const uint32_t f0 = s.out_offset + level * F;  // multiple of F
            half_t h[F];
#pragma unroll
            for (int f = 0; f < F; ++f) h[f] = valid ? (half_t)v[f] : (half_t)0.0f;
            if constexpr (F == 1) {
                feat[feat_pair_offset(q >> 4, f0 >> 1, kp2, q & 15) + (f0 & 1)] = h[0];
            } else {
#pragma unroll
                for (int j = 0; j < F / 2; ++j) *pair_ptr(feat, q, f0 + 2 * j, kp2) = pack2(h[2 * j], h[2 * j + 1]);
            }
        }
    }
}

struct MlpArgs16 {
    uint32_t N, n_output, n_hidden_layers, k_padded;
    int32_t act, out_act;
    uint32_t packed_halves;
};

// TT tiles of 16 queries per wave iteration: each weight fragment read from
// LDS feeds TT MFMAs (LDS bandwidth for the weights bounds the single-tile loop).
// ACT: the hidden activation when fixed at compile time (ReLU), or -1 to read it
// from the arguments.  KC0 > 0: the input width in 32-feature chunks, fixed at
// compile time, and each wave loads the feature fragments of its next group
// while it computes the current one (the loads are then in flight for a whole
// group's MFMAs); KC0 = 0: width from the arguments, loaded as used.
template <int W, int TT, int ACT, int KC0>
__global__ __launch_bounds__(kBlock) void mlp_fwd_kernel(MlpArgs16 a, const half_t* __restrict__ packed,
                                                         const half_t* __restrict__ feat, float* __restrict__ out) {
    constexpr int MT = W / 16, KC = W / 32;
    constexpr int PF = KC0 > 0 ? KC0 : 1;
    const int act = ACT >= 0 ? ACT : a.act;
    extern __shared__ float4 smem4[];
    half_t* sw = reinterpret_cast<half_t*>(smem4);
    for (uint32_t k = threadIdx.x; k < a.packed_halves / 8; k += kBlock)
        reinterpret_cast<float4*>(sw)[k] = reinterpret_cast<const float4*>(packed)[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const uint32_t kc0 = KC0 > 0 ? (uint32_t)KC0 : a.k_padded / 32, kp2 = a.k_padded / 2;
    const uint32_t n_tiles = (a.N + 15) / 16;
    const uint32_t n_groups = (n_tiles + TT - 1) / TT;
    const uint32_t n_waves = gridDim.x * (kBlock / 64);
    const uint32_t grp0 = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    half8 nb[TT][PF];  // KC0 > 0: fragments of the wave's next group
    auto load_group = [&](uint32_t gi) {
#pragma unroll
        for (int t = 0; t < TT; ++t) {
            const uint32_t tile = min(gi * TT + t, n_tiles - 1);  // a clamped duplicate is not stored
#pragma unroll
            for (int c = 0; c < PF; ++c) nb[t][c] = load_feat_frag(feat, tile, kp2, c, g, n);
        }
    };
    if (KC0 > 0 && grp0 < n_groups) load_group(grp0);
    for (uint32_t grp = grp0; grp < n_groups; grp += n_waves) {
        floatx4 acc[TT][MT];
#pragma unroll
        for (int t = 0; t < TT; ++t)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) acc[t][mt] = floatx4{0.f, 0.f, 0.f, 0.f};
        if constexpr (KC0 > 0) {
            half8 b[TT][PF];
#pragma unroll
            for (int t = 0; t < TT; ++t)
#pragma unroll
                for (int c = 0; c < PF; ++c) b[t][c] = nb[t][c];
            if (grp + n_waves < n_groups) load_group(grp + n_waves);
#pragma unroll
            for (int c = 0; c < PF; ++c)
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    const half8 w = *reinterpret_cast<const half8*>(sw + ((size_t)(mt * PF + c) * 64 + lane) * 8);
#pragma unroll
                    for (int t = 0; t < TT; ++t) acc[t][mt] = mfma16(w, b[t][c], acc[t][mt]);
                }
        } else {
            for (uint32_t c = 0; c < kc0; ++c) {
                half8 b[TT];
#pragma unroll
                for (int t = 0; t < TT; ++t) {
                    const uint32_t tile = min(grp * TT + t, n_tiles - 1);
                    b[t] = load_feat_frag(feat, tile, kp2, c, g, n);
                }
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    const half8 w = *reinterpret_cast<const half8*>(sw + ((size_t)(mt * kc0 + c) * 64 + lane) * 8);
#pragma unroll
                    for (int t = 0; t < TT; ++t) acc[t][mt] = mfma16(w, b[t], acc[t][mt]);
                }
            }
        }
        const half_t* frag = sw + (size_t)MT * kc0 * 512;
        half8 hb[TT][KC];
#pragma unroll
        for (int t = 0; t < TT; ++t)
#pragma unroll
            for (int c = 0; c < KC; ++c) hb[t][c] = pack_act(acc[t][2 * c], acc[t][2 * c + 1], act);
        for (uint32_t layer = 1; layer < a.n_hidden_layers; ++layer) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                floatx4 s[TT];
#pragma unroll
                for (int t = 0; t < TT; ++t) s[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int c = 0; c < KC; ++c) {
                    const half8 w = *reinterpret_cast<const half8*>(frag + ((size_t)(mt * KC + c) * 64 + lane) * 8);
#pragma unroll
                    for (int t = 0; t < TT; ++t) s[t] = mfma16(w, hb[t][c], s[t]);
                }
#pragma unroll
                for (int t = 0; t < TT; ++t) acc[t][mt] = s[t];
            }
            frag += (size_t)MT * KC * 512;
#pragma unroll
            for (int t = 0; t < TT; ++t)
#pragma unroll
                for (int c = 0; c < KC; ++c) hb[t][c] = pack_act(acc[t][2 * c], acc[t][2 * c + 1], act);
        }
        floatx4 o[TT];
#pragma unroll
        for (int t = 0; t < TT; ++t) o[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            const half8 w = *reinterpret_cast<const half8*>(frag + ((size_t)c * 64 + lane) * 8);
#pragma unroll
            for (int t = 0; t < TT; ++t) o[t] = mfma16(w, hb[t][c], o[t]);
        }
#pragma unroll
        for (int t = 0; t < TT; ++t) {
            const uint32_t tile = grp * TT + t;
            const uint32_t q = tile * 16 + n;
            if (tile < n_tiles && q < a.N) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t oi = 4 * g + i;
                    if (oi < a.n_output) out[(size_t)q * a.n_output + oi] = act_fwd(a.out_act, o[t][i]);
                }
            }
        }
    }
}

// ReLU networks with inputs up to 128 features take the prefetching kernels
template <int W, int TT, typename Launch>
void launch_fwd(const MlpArgs16& a, Launch&& launch) {
    if (a.act != ACT_RELU) return launch(mlp_fwd_kernel<W, TT, -1, 0>, TT);
    switch (a.k_padded / 32) {
        case 1: return launch(mlp_fwd_kernel<W, TT, ACT_RELU, 1>, TT);
        case 2: return launch(mlp_fwd_kernel<W, TT, ACT_RELU, 2>, TT);
        case 3: return launch(mlp_fwd_kernel<W, TT, ACT_RELU, 3>, TT);
        case 4: return launch(mlp_fwd_kernel<W, TT, ACT_RELU, 4>, TT);
        default: return launch(mlp_fwd_kernel<W, TT, ACT_RELU, 0>, TT);
    }
}

__global__ void to_half_kernel(const float* __restrict__ src, half_t* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (half_t)src[i];
}

// Splits the grid levels between the LDS pass (levels whose tables still fit,
// in order) and the rows of the per-level pass (the rest).
void make_level_maps(const Layout& L, LevelMap& m, LdsLevels& d) {
    m = LevelMap{};
    d = LdsLevels{};
    for (size_t si = 0; si < L.segs.size(); ++si) {
        const Segment& s = L.segs[si];
        if (s.type != SEG_GRID) continue;
        for (uint32_t l = 0; l < s.n_levels; ++l) {
            const uint32_t halves = (s.levels.offset[l + 1] - s.levels.offset[l]) * s.F;
            if (d.n < (uint32_t)kMaxLdsLevels && (size_t)(d.halves + halves) * sizeof(half_t) <= kLdsTableBytes) {
                d.seg[d.n] = (uint8_t)si;
                d.level[d.n] = (uint8_t)l;
                d.at[d.n] = d.halves;
                d.halves += halves;
                ++d.n;
                continue;
            }
            if (m.n >= (uint32_t)kMaxLevelRows) fail(NLOD_ERR_UNSUPPORTED, "more than 64 grid levels in one model");
            m.seg[m.n] = (uint8_t)si;
            m.level[m.n] = (uint8_t)l;
            ++m.n;
        }
    }
}

}  // namespace

size_t mfma_fused_smem_bytes(const NetShape& net) { return mfma_packed_bytes(net); }

bool mfma_supported(const Layout& L) {
    if (L.net.width != 32 && L.net.width != 64 && L.net.width != 128) return false;
    if (L.segs.size() > (size_t)kMaxSegs) return false;
    if (L.net.n_output > 16) return false;
    if (k_padded_of(L.net) > 8 * kMaxGroups) return false;
    if (mfma_fused_smem_bytes(L.net) > 160 * 1024) return false;
    uint32_t F = 0, levels = 0;
    for (const Segment& s : L.segs) {
        if (s.type != SEG_GRID) continue;
        if (F && s.F != F) return false;  // one feature count per model
        F = s.F;
        levels += s.n_levels;
    }
    return levels <= (uint32_t)kMaxLevelRows;
}

size_t mfma_feature_bytes(const Layout& L, uint32_t N) {
    return sizeof(half_t) * (size_t)div_round_up(N, 16) * 16 * k_padded_of(L.net);
}

void mfma_to_half(hipStream_t stream, const float* src, void* dst, size_t n) {
    if (n == 0) return;
    const uint32_t blocks = (uint32_t)std::min<size_t>(div_round_up(n, 256), 8192);
    hipLaunchKernelGGL(to_half_kernel, dim3(blocks), dim3(256), 0, stream, src, static_cast<half_t*>(dst), n);
    NLOD_LAUNCH_CHECK();
}

size_t mfma_pos_scratch_bytes(const Layout& L, uint32_t N) {
    LevelMap map;
    LdsLevels lds;
    make_level_maps(L, map, lds);
    bool any = false, same = true;
    uint32_t off = 0;
    for (const Segment& s : L.segs) {
        if (s.type != SEG_GRID) continue;
        if (any && s.in_offset != off) same = false;
        off = s.in_offset;
        any = true;
    }
    if (!lds.n || !any || !same || map.n == 0) return 0;
    return sizeof(float) * 3 * (size_t)(div_round_up(N, 16) * 16);
}

void mfma_encode(hipStream_t stream, const Layout& L, const void* d_program, const void* table16, const float* in,
                 uint32_t N, void* feat, void* pos_scratch) {
    if (N == 0) return;
    LevelMap map;
    LdsLevels lds;
    make_level_maps(L, map, lds);
    const uint32_t N16 = div_round_up(N, 16) * 16;
    // positions as [3][N16], written by the LDS pass and read by the level
    // rows, when every grid segment reads the same 3 inputs
    float* xsoa = nullptr;
    uint32_t soa_off = 0;
    if (pos_scratch && mfma_pos_scratch_bytes(L, N) > 0) {
        xsoa = static_cast<float*>(pos_scratch);
        for (const Segment& s : L.segs)
            if (s.type == SEG_GRID) soa_off = s.in_offset;
    }
    if (lds.n) {
        const size_t smem = (size_t)lds.halves * sizeof(half_t);
        const uint32_t blocks = std::min<uint32_t>(div_round_up(N16, kLdsBlock), (uint32_t)cu_count() * 2);
        dispatch_fused(L, [&](auto Wc, auto Fc, auto Dc) {
            constexpr int F = decltype(Fc)::value;
            NLOD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(encode_lds_kernel<F>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsTableBytes));
            hipLaunchKernelGGL(encode_lds_kernel<F>, dim3(blocks), dim3(kLdsBlock), smem, stream, N, N16, lds,
                               static_cast<const EncProgram*>(d_program), static_cast<const half_t*>(table16), in,
                               static_cast<half_t*>(feat), xsoa, soa_off);
        });
        NLOD_LAUNCH_CHECK();
    }
    const dim3 grid(div_round_up(2 * N16, kBlock), map.n + 1);  // grid rows take two lanes per query
    dispatch_fused(L, [&](auto Wc, auto Fc, auto Dc) {
        constexpr int F = decltype(Fc)::value, D = decltype(Dc)::value;
        hipLaunchKernelGGL((encode_kernel<F, D, half_t>), grid, dim3(kBlock), 0, stream, N, N16, map,
                           static_cast<const EncProgram*>(d_program), static_cast<const half_t*>(table16), in,
                           static_cast<half_t*>(feat), (const float*)xsoa);
    });
    NLOD_LAUNCH_CHECK();
}

void mfma_mlp_forward(hipStream_t stream, const Layout& L, const void* packed, const void* feat, uint32_t N,
                      float* out) {
    if (N == 0) return;
    MlpArgs16 a{};
    a.N = N;
    a.n_output = L.net.n_output;
    a.n_hidden_layers = L.net.n_hidden_layers;
    a.k_padded = k_padded_of(L.net);
    a.act = L.net.activation;
    a.out_act = L.net.output_activation;
    a.packed_halves = (uint32_t)packed_halves(L.net);
    const size_t smem = mfma_fused_smem_bytes(L.net);
    const uint32_t n_tiles = div_round_up(N, 16);
    auto launch = [&](auto kern, uint32_t tt) {
        const int occ = occupancy_blocks((const void*)kern, kBlock, smem);
        const uint32_t groups = div_round_up(n_tiles, tt);
        const uint32_t blocks = std::min<uint32_t>(div_round_up(groups, kBlock / 64), (uint32_t)(cu_count() * occ));
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), smem, stream, a, static_cast<const half_t*>(packed),
                           static_cast<const half_t*>(feat), out);
    };
    switch (L.net.width) {
        case 32: launch_fwd<32, 2>(a, launch); break;
        case 64: launch_fwd<64, 2>(a, launch); break;
        case 128: launch_fwd<128, 1>(a, launch); break;
        default: fail(NLOD_ERR_UNSUPPORTED, "FP16 MFMA kernels support widths 32, 64 and 128");
    }
    NLOD_LAUNCH_CHECK();
}

}  // namespace nlod
