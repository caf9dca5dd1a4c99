// Shared pieces of the FP16-MFMA kernels: the device-side encoding program,
// the per-lane feature-window encoder, the packed weight-fragment layout and
// the register-resident MLP tail.  See mlp_mfma.hip for the operand layout.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "device_common.h"
#include "layout.h"
#include "nlod_common.h"
#include "sh_basis.h"

namespace nlod {

constexpr int kMaxSegs = 8;

// Encoding description read by the fused kernels (uniform, from device memory).
struct DevSegment {
    int32_t type;
    uint32_t in_offset, out_offset, n_out, n_pad, degree;
    float id_scale, id_offset;
    uint64_t param_offset;
    GridLevels lv;
};
constexpr int kMaxGroups = 64;  // 8-feature groups of the padded encoding (k_padded <= 512)
struct EncProgram {
    uint32_t n_segs, n_input_dims, enc_width, k_padded;
    DevSegment segs[kMaxSegs];
    // bit j of grid_rows[g] set: encoded row 8g + j is a grid feature (written by the per-level pass)
    uint8_t grid_rows[kMaxGroups];
};

inline uint32_t k_padded_of(const NetShape& n) { return (n.n_input + 31) / 32 * 32; }

// FP16 encoded features of a batch: [N/16 tiles][K/2 feature pairs][16 queries][2 halves].
// A grid level with F = 2 fills one contiguous 64-byte row of a tile, so the
// per-level encoding pass writes whole cache lines.  Offset in halves:
__host__ __device__ inline size_t feat_pair_offset(uint32_t tile, uint32_t pair, uint32_t kp2, uint32_t n) {
    return (((size_t)tile * kp2 + pair) * 16 + n) * 2;
}

// B fragment of k-chunk c for lane group g and query n of a tile: features
// 32c + 8g .. 32c + 8g + 7, i.e. pairs 16c + 4g .. 16c + 4g + 3 (4 dwords, 64 B apart).
__device__ __forceinline__ half8 load_feat_frag(const half_t* __restrict__ feat, uint32_t tile, uint32_t kp2,
                                                uint32_t c, uint32_t g, uint32_t n) {
    const uint32_t* p = reinterpret_cast<const uint32_t*>(feat) + ((size_t)tile * kp2 + 16 * c + 4 * g) * 16 + n;
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 w = {p[0], p[16], p[32], p[48]};
    return __builtin_bit_cast(half8, w);
}

// Packed FP16 weight fragments: per matrix, per 16-row tile, per 32-column
// chunk, 64 lanes x 8 halves (layer 0 in natural column order, later layers
// with columns permuted to the accumulator layout).
inline size_t frag_count(uint32_t rows, uint32_t cols_padded) { return (size_t)(rows / 16) * (cols_padded / 32); }

inline size_t packed_halves(const NetShape& n) {
    const uint32_t W = n.width;
    const size_t frags = frag_count(W, k_padded_of(n)) + (size_t)(n.n_hidden_layers - 1) * frag_count(W, W) +
                         frag_count(16, W);
    return frags * 64 * 8;
}

// k-slot (g, j) of chunk c of a layer fed by the previous layer's accumulators
// -> hidden unit: slots 0-3 come from accumulator tile 2c, 4-7 from tile 2c+1.
__host__ __device__ inline uint32_t permuted_unit(uint32_t c, uint32_t g, uint32_t j) {
    return 32u * c + (j < 4 ? 4u * g + j : 16u + 4u * g + (j - 4u));
}

// Packed transposed weights for the backward products (same fragment order as
// the forward packing: per matrix, per 16-row tile, per 32-column chunk).
//   output layer: W_out^T [W][32]  (columns = outputs, rows >= 16 zero)
//   hidden l (L-1 .. 1): W_l^T [W][W]
//   input layer: W_0^T [Kp][W]  (rows >= K zero)
// Columns are permuted to the accumulator layout of the delta they multiply.
inline size_t packed_bwd_halves(const NetShape& n) {
    const uint32_t W = n.width, Kp = k_padded_of(n);
    const size_t frags = frag_count(W, 32) + (size_t)(n.n_hidden_layers - 1) * frag_count(W, W) + frag_count(Kp, W);
    return frags * 512;
}

// Element e of the forward packing of the weight vector w (W wide, K inputs
// padded to Kp, L hidden layers).
__device__ inline float packed_fwd_value(uint32_t W, uint32_t K, uint32_t Kp, uint32_t L, const float* __restrict__ w,
                                         uint32_t e) {
    const uint32_t j = e & 7u, lane = (e >> 3) & 63u, frag = e >> 9;
    const uint32_t g = lane >> 4, r16 = lane & 15u;
    const uint32_t f0 = (W / 16) * (Kp / 32);
    const uint32_t fh = (W / 16) * (W / 32);
    if (frag < f0) {  // input layer: natural column order, zero beyond K
        const uint32_t mt = frag / (Kp / 32), c = frag % (Kp / 32);
        const uint32_t row = 16 * mt + r16, col = 32 * c + 8 * g + j;
        return col < K ? w[(size_t)row * K + col] : 0.0f;
    }
    if (frag < f0 + (L - 1) * fh) {  // hidden matrices: permuted columns
        const uint32_t f = frag - f0, layer = 1 + f / fh, ff = f % fh;
        const uint32_t mt = ff / (W / 32), c = ff % (W / 32);
        const uint32_t row = 16 * mt + r16, col = permuted_unit(c, g, j);
        return w[(size_t)W * K + (size_t)(layer - 1) * W * W + (size_t)row * W + col];
    }
    // output matrix (16 padded rows): permuted columns
    const uint32_t c = frag - f0 - (L - 1) * fh;
    const uint32_t row = r16, col = permuted_unit(c, g, j);
    return w[(size_t)W * K + (size_t)(L - 1) * W * W + (size_t)row * W + col];
}

// Element e of the backward packing of the weight vector w.
__device__ inline float packed_bwd_value(uint32_t W, uint32_t K, uint32_t L, const float* __restrict__ w, uint32_t e) {
    const uint32_t j = e & 7u, lane = (e >> 3) & 63u, frag = e >> 9;
    const uint32_t g = lane >> 4, r16 = lane & 15u;
    const uint32_t f_out = W / 16;  // output-layer fragments (1 chunk each)
    const uint32_t f_hid = (W / 16) * (W / 32);
    const size_t off_out = (size_t)W * K + (size_t)(L - 1) * W * W;
    if (frag < f_out) {
        const uint32_t mt = frag;
        const uint32_t row = 16 * mt + r16;           // input unit of W_out
        const uint32_t col = permuted_unit(0, g, j);  // output index
        return col < 16 ? w[off_out + (size_t)col * W + row] : 0.0f;
    }
    if (frag < f_out + (L - 1) * f_hid) {
        const uint32_t f = frag - f_out, idx = f / f_hid, ff = f % f_hid;
        const uint32_t layer = L - 1 - idx;  // hidden matrices in backward order
        const uint32_t mt = ff / (W / 32), c = ff % (W / 32);
        const uint32_t row = 16 * mt + r16, col = permuted_unit(c, g, j);
        return w[(size_t)W * K + (size_t)(layer - 1) * W * W + (size_t)col * W + row];
    }
    const uint32_t f = frag - f_out - (L - 1) * f_hid;
    const uint32_t mt = f / (W / 32), c = f % (W / 32);
    const uint32_t row = 16 * mt + r16, col = permuted_unit(c, g, j);
    return row < K ? w[(size_t)col * K + row] : 0.0f;
}

// Table entry of F features (FP32 or FP16 storage) as floats.
template <int F, typename TT>
__device__ __forceinline__ void load_entry(const TT* __restrict__ p, float* out) {
    if constexpr (std::is_same<TT, float>::value) {
        if constexpr (F == 1) {
            out[0] = p[0];
        } else if constexpr (F == 2) {
            const float2 t2 = *reinterpret_cast<const float2*>(p);
            out[0] = t2.x;
            out[1] = t2.y;
        } else {
#pragma unroll
            for (int f = 0; f < F; f += 4) {
                const float4 t4 = *reinterpret_cast<const float4*>(p + f);
                out[f] = t4.x;
                out[f + 1] = t4.y;
                out[f + 2] = t4.z;
                out[f + 3] = t4.w;
            }
        }
    } else {
        if constexpr (F == 1) {
            out[0] = (float)p[0];
        } else if constexpr (F == 2) {
            typedef _Float16 h2 __attribute__((ext_vector_type(2)));
            const h2 t = *reinterpret_cast<const h2*>(p);
            out[0] = (float)t[0];
            out[1] = (float)t[1];
        } else {
#pragma unroll
            for (int f = 0; f < F; f += 4) {
                const half4 t = *reinterpret_cast<const half4*>(p + f);
                out[f] = (float)t[0];
                out[f + 1] = (float)t[1];
                out[f + 2] = (float)t[2];
                out[f + 3] = (float)t[3];
            }
        }
    }
}

// One level of a grid for one query: F features interpolated in FP32, from the
// level's own entries at tab (global memory, or a copy in LDS).
template <int F, typename TT = float>
__device__ __forceinline__ void grid_level_at(const GridLevels& lv, uint32_t level, float x0, float x1, float x2,
                                              const TT* __restrict__ tab, float* v) {
    // no FMA contraction: the encoding rounds identically in every kernel it is inlined into
#pragma clang fp contract(off)
    const float scale = lv.scale[level];
    const uint32_t res = lv.res[level];
    const uint32_t hs = lv.offset[level + 1] - lv.offset[level];
    float fr[3];
    uint32_t cell[3];
    grid_pos(x0, scale, lv.interp, fr[0], cell[0]);
    grid_pos(x1, scale, lv.interp, fr[1], cell[1]);
    grid_pos(x2, scale, lv.interp, fr[2], cell[2]);
#pragma unroll
    for (int f = 0; f < F; ++f) v[f] = 0.0f;
    if (lv.interp == INTERP_NEAREST) {
        const uint32_t idx = grid_index(lv.grid_type, hs, res, cell[0], cell[1], cell[2]);
        load_entry<F, TT>(tab + (size_t)idx * F, v);
        return;
    }
    float vals[8][F];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const uint32_t idx = grid_index(lv.grid_type, hs, res, cell[0] + (c & 1), cell[1] + ((c >> 1) & 1),
                                        cell[2] + ((c >> 2) & 1));
        load_entry<F, TT>(tab + (size_t)idx * F, vals[c]);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float w = 1.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) w *= ((c >> d) & 1) ? fr[d] : 1.0f - fr[d];
#pragma unroll
        for (int f = 0; f < F; ++f) v[f] += w * vals[c][f];
    }
}

// One level of a grid for one query from the whole table.
template <int F, typename TT = float>
__device__ __forceinline__ void grid_level(const GridLevels& lv, uint32_t level, float x0, float x1, float x2,
                                           const TT* __restrict__ table, float* v) {
    grid_level_at<F, TT>(lv, level, x0, x1, x2, table + (size_t)lv.offset[level] * F, v);
}

// Entry e (0 <= e < 2 BW / F) of a block of BW 32-bit words of FP16 entries
// with F features.
template <int F, int BW>
__device__ __forceinline__ void entry_of_block(const uint32_t (&w)[BW], uint32_t e, float* out) {
    if constexpr (2 * BW == F) {  // one entry per block
#pragma unroll
        for (int k = 0; k < BW; ++k) {
            out[2 * k] = (float)__builtin_bit_cast(half_t, (uint16_t)(w[k] & 0xffffu));
            out[2 * k + 1] = (float)__builtin_bit_cast(half_t, (uint16_t)(w[k] >> 16));
        }
    } else {
        constexpr uint32_t WORDS = F == 1 ? 1 : F / 2;  // 32-bit words per entry (F = 1: half a word)
        const uint32_t word0 = F == 1 ? e / 2 : e * WORDS;
        uint32_t sel[WORDS];
#pragma unroll
        for (uint32_t k = 0; k < WORDS; ++k) {
            const uint32_t wi = word0 + k;
            uint32_t x = w[0];
#pragma unroll
            for (int j = 1; j < BW; ++j) x = wi == (uint32_t)j ? w[j] : x;
            sel[k] = x;
        }
        if constexpr (F == 1) {
            const uint32_t bits = (e & 1) ? (sel[0] >> 16) : (sel[0] & 0xffffu);
            out[0] = (float)__builtin_bit_cast(half_t, (uint16_t)bits);
        } else {
#pragma unroll
            for (uint32_t k = 0; k < WORDS; ++k) {
                out[2 * k] = (float)__builtin_bit_cast(half_t, (uint16_t)(sel[k] & 0xffffu));
                out[2 * k + 1] = (float)__builtin_bit_cast(half_t, (uint16_t)(sel[k] >> 16));
            }
        }
    }
}

// Loads BW 32-bit words (BW = 2 or 4) from a BW-word-aligned address.
template <int BW>
__device__ __forceinline__ void load_block(const half_t* __restrict__ p, uint32_t (&w)[BW]) {
    if constexpr (BW == 4) {
        const uint4 t = *reinterpret_cast<const uint4*>(p);
        w[0] = t.x;
        w[1] = t.y;
        w[2] = t.z;
        w[3] = t.w;
    } else {
        const uint2 t = *reinterpret_cast<const uint2*>(p);
        w[0] = t.x;
        w[1] = t.y;
    }
}

// One level of a grid for one query from an FP16 table, F features in FP32.
// The two corners of an x-edge usually share an aligned block of the table
// (dense levels: adjacent entries; coherent prime hash: x enters with prime 1,
// so the entries differ only in the low bits x0 ^ (x0 + 1)): one block load
// then serves both, and only the other cases issue a second load.  Blocks are
// 8 bytes (F <= 4) or 16 bytes (F = 8): an 8-byte block holds both corners of
// an edge half the time for F = 2, and takes half the registers of a 16-byte
// one, which matters to the kernels that keep a wave's gathers of 4 levels in
// flight at once.  Same interpolation arithmetic and order as grid_level.
template <int F>
__device__ __forceinline__ void grid_level_h16(const GridLevels& lv, uint32_t level, float x0, float x1, float x2,
                                               const half_t* __restrict__ table, float* v) {
    // no FMA contraction: the encoding rounds identically in every kernel it is inlined into
#pragma clang fp contract(off)
    constexpr int BW = F == 8 ? 4 : 2;                // 32-bit words per block
    constexpr uint32_t PER_BLOCK = 2 * BW / F;        // entries per block
    const float scale = lv.scale[level];
    const uint32_t res = lv.res[level];
    const uint32_t off = lv.offset[level];
    const uint32_t hs = lv.offset[level + 1] - off;
    const half_t* tab = table + (size_t)off * F;
    float fr[3];
    uint32_t cell[3];
    grid_pos(x0, scale, lv.interp, fr[0], cell[0]);
    grid_pos(x1, scale, lv.interp, fr[1], cell[1]);
    grid_pos(x2, scale, lv.interp, fr[2], cell[2]);
#pragma unroll
    for (int f = 0; f < F; ++f) v[f] = 0.0f;
    if (lv.interp == INTERP_NEAREST) {
        const uint32_t idx = grid_index(lv.grid_type, hs, res, cell[0], cell[1], cell[2]);
        load_entry<F, half_t>(tab + (size_t)idx * F, v);
        return;
    }
    // p = (dy, dz); corners c = 2p (x0) and 2p + 1 (x0 + 1).  All block loads are
    // issued before any is used, so a lane keeps its up to 8 gathers in flight
    // together instead of waiting for each in turn.
    uint32_t i0[4], i1[4];
    {
        const LevelIndexer li = level_indexer(lv.grid_type, hs, res);
        const uint32_t y0 = cell[1] * li.my, z0 = cell[2] * li.mz;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const uint32_t ym = (p & 1) ? y0 + li.my : y0, zm = ((p >> 1) & 1) ? z0 + li.mz : z0;
            i0[p] = level_index(li, cell[0], ym, zm);
            i1[p] = level_index(li, cell[0] + 1, ym, zm);
        }
    }
    uint32_t blk0[4][BW], blk1[4][BW];
#pragma unroll
    for (int p = 0; p < 4; ++p) load_block<BW>(tab + (size_t)(i0[p] / PER_BLOCK) * PER_BLOCK * F, blk0[p]);
    bool shared[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        // the x0 + 1 corner may share x0's block; otherwise it needs its own.
        // The load is issued either way, without a branch (a branch makes the
        // compiler wait for the loads before it): lanes whose corner is shared
        // read the level's first block, one line that the whole wave shares
        // and that stays in the vector L1, so it costs no L2 request per lane.
        shared[p] = i0[p] / PER_BLOCK == i1[p] / PER_BLOCK;
        const size_t b1 = shared[p] ? 0 : (size_t)(i1[p] / PER_BLOCK) * PER_BLOCK * F;
        load_block<BW>(tab + b1, blk1[p]);
    }
    float vals[8][F];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        entry_of_block<F, BW>(blk0[p], i0[p] % PER_BLOCK, vals[2 * p]);
        uint32_t b1w[BW];
#pragma unroll
        for (int j = 0; j < BW; ++j) b1w[j] = shared[p] ? blk0[p][j] : blk1[p][j];
        entry_of_block<F, BW>(b1w, i1[p] % PER_BLOCK, vals[2 * p + 1]);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float w = 1.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) w *= ((c >> d) & 1) ? fr[d] : 1.0f - fr[d];
#pragma unroll
        for (int f = 0; f < F; ++f) v[f] += w * vals[c][f];
    }
}

// One level of a grid for one query from an FP16 table, on two lanes: an even
// lane and the odd lane after it (s = lane parity) share the query.  Lane s
// gathers the x0 + s corner of each of the 4 (y, z) rows, so within one load
// instruction the two corners of an x-edge come from adjacent lanes, and since
// they almost always lie in one cache line the pair costs one L2 request
// rather than two.  The lanes then swap their gathers (DPP quad_perm
// [1, 0, 3, 2]) and lane s interpolates the features of parity s: v[k] is
// feature 2k + s (F = 1: both lanes compute feature 0).  Per feature, the same
// arithmetic and order as grid_level.  Both lanes of a pair must be active.
template <int F>
__device__ __forceinline__ void grid_level_h16_pair(const GridLevels& lv, uint32_t level, float x0, float x1,
                                                    float x2, const half_t* __restrict__ table, uint32_t s,
                                                    float* v) {
#pragma clang fp contract(off)
    constexpr int WORDS = F == 1 ? 1 : F / 2;  // 32-bit words per entry (F = 1: the half in the low bits)
    const float scale = lv.scale[level];
    const uint32_t res = lv.res[level];
    const uint32_t off = lv.offset[level];
    const uint32_t hs = lv.offset[level + 1] - off;
    const half_t* tab = table + (size_t)off * F;
    float fr[3];
    uint32_t cell[3];
    grid_pos(x0, scale, lv.interp, fr[0], cell[0]);
    grid_pos(x1, scale, lv.interp, fr[1], cell[1]);
    grid_pos(x2, scale, lv.interp, fr[2], cell[2]);
    auto load_words = [&](uint32_t idx, uint32_t* w) {
        const half_t* e = tab + (size_t)idx * F;
        if constexpr (F == 1) {
            w[0] = __builtin_bit_cast(uint16_t, e[0]);
        } else if constexpr (F == 2) {
            w[0] = *reinterpret_cast<const uint32_t*>(e);
        } else if constexpr (F == 4) {
            const uint2 t = *reinterpret_cast<const uint2*>(e);
            w[0] = t.x;
            w[1] = t.y;
        } else {
            const uint4 t = *reinterpret_cast<const uint4*>(e);
            w[0] = t.x;
            w[1] = t.y;
            w[2] = t.z;
            w[3] = t.w;
        }
    };
    // feature 2k + s (F = 1: feature 0) of an entry held as words
    auto feature = [&](const uint32_t* w, int k) -> float {
        const uint32_t bits = F == 1 ? w[0] : (w[k] >> (16 * s)) & 0xffffu;
        return (float)__builtin_bit_cast(half_t, (uint16_t)bits);
    };
#pragma unroll
    for (int k = 0; k < WORDS; ++k) v[k] = 0.0f;
    if (lv.interp == INTERP_NEAREST) {
        uint32_t w[WORDS];
        load_words(grid_index(lv.grid_type, hs, res, cell[0], cell[1], cell[2]), w);
#pragma unroll
        for (int k = 0; k < WORDS; ++k) v[k] = feature(w, k);
        return;
    }
    uint32_t mine[4][WORDS], other[4][WORDS];
    {
        const LevelIndexer li = level_indexer(lv.grid_type, hs, res);
        const uint32_t y0 = cell[1] * li.my, z0 = cell[2] * li.mz;
#pragma unroll
        for (int p = 0; p < 4; ++p)
            load_words(level_index(li, cell[0] + s, (p & 1) ? y0 + li.my : y0, ((p >> 1) & 1) ? z0 + li.mz : z0),
                       mine[p]);
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int k = 0; k < WORDS; ++k)
            other[p][k] = (uint32_t)__builtin_amdgcn_mov_dpp((int)mine[p][k], 0xB1, 0xF, 0xF, false);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float w = 1.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) w *= ((c >> d) & 1) ? fr[d] : 1.0f - fr[d];
        const uint32_t* e = ((uint32_t)(c & 1) == s) ? mine[c >> 1] : other[c >> 1];
#pragma unroll
        for (int k = 0; k < WORDS; ++k) v[k] += w * feature(e, k);
    }
}

// win[j] = v for a runtime window slot j, via selects (no dynamic register indexing).
__device__ __forceinline__ void place(float* win, int j, float v) {
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) win[jj] = (jj == j) ? v : win[jj];
}

// Encoded features [w0, w0 + 8) of one query in FP32 (0 beyond the encoded width;
// grid rows left at 0 when SKIP_GRIDS).
template <int F, int MAXD, typename TT = float, bool SKIP_GRIDS = false>
__device__ __forceinline__ void encode_window_f32(const EncProgram& P, const TT* __restrict__ params,
                                                  const float* __restrict__ q, uint32_t w0, float* win) {
    // no FMA contraction: the encoding rounds identically in every kernel it is inlined into
#pragma clang fp contract(off)
#pragma unroll
    for (int j = 0; j < 8; ++j) win[j] = 0.0f;
    const uint32_t w1 = w0 + 8;
    for (uint32_t si = 0; si < P.n_segs; ++si) {
        const DevSegment& s = P.segs[si];
        const uint32_t r0 = s.out_offset, r1 = s.out_offset + s.n_out + s.n_pad;
        if (r1 <= w0 || r0 >= w1) continue;
        const float* x = q + s.in_offset;
        if (s.type == SEG_GRID) {
            if (SKIP_GRIDS) continue;
            // a grid segment starts at a multiple of F, so levels tile the window
            const int l_begin = w0 > r0 ? (int)((w0 - r0) / F) : 0;
            const int n_levels = (int)s.lv.n_levels;
            const float x0 = x[0], x1 = x[1], x2 = x[2];
            const TT* table = params + s.param_offset;
            // rows of level l_begin + t start at row0 + t F; row0 = 0 when the
            // window starts at or after the segment (always for a segment at a
            // window boundary), so the placement is by constant slot
            const int row0 = (int)r0 + l_begin * F - (int)w0;
#pragma unroll
            for (int t = 0; t < 8 / F; ++t) {
                const int l = l_begin + t;
                const int row = row0 + t * F;
                // gather unconditionally (clamped level: indices stay inside it), place if valid
                float v[F];
                if constexpr (std::is_same<TT, half_t>::value)
                    grid_level_h16<F>(s.lv, (uint32_t)min(l, n_levels - 1), x0, x1, x2, table, v);
                else
                    grid_level<F, TT>(s.lv, (uint32_t)min(l, n_levels - 1), x0, x1, x2, table, v);
                const bool ok = l < n_levels && row >= 0 && row < 8;
                if (row0 == 0) {
#pragma unroll
                    for (int f = 0; f < F; ++f) win[t * F + f] = l < n_levels ? v[f] : win[t * F + f];
                } else {
#pragma unroll
                    for (int f = 0; f < F; ++f) place(win, ok ? row + f : -1, v[f]);
                }
            }
        } else if (s.type == SEG_SH) {
            const uint32_t sh0 = r0 + s.n_pad;  // SH pad rows (1.0) come first
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t row = w0 + j;
                if (row >= r0 && row < sh0) win[j] = 1.0f;
            }
            if (sh0 + s.degree * s.degree > w0 && sh0 < w1) {
                const float vx = x[0] * 2.f - 1.f, vy = x[1] * 2.f - 1.f, vz = x[2] * 2.f - 1.f;
                float X[MAXD], Y[MAXD], Z[MAXD];
                X[0] = Y[0] = Z[0] = 1.0f;
#pragma unroll
                for (int k = 1; k < MAXD; ++k) {
                    X[k] = X[k - 1] * vx;
                    Y[k] = Y[k - 1] * vy;
                    Z[k] = Z[k - 1] * vz;
                }
                const int base = (int)sh0 - (int)w0;
                if ((base & 7) == 0) {
                    // coefficient k lands in slot k & 7 of window -base / 8: one select each
                    const int wk = -base >> 3;
                    sh_basis<MAXD>(s.degree, X, Y, Z, [&](int k, float v) {
                        if ((k >> 3) == wk) win[k & 7] = v;
                    });
                } else {
                    sh_basis<MAXD>(s.degree, X, Y, Z, [&](int k, float v) { place(win, base + k, v); });
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t row = w0 + j;
                if (row >= r0 && row < r0 + s.n_out) win[j] = x[row - r0] * s.id_scale + s.id_offset;
                else if (row >= r0 + s.n_out && row < r1) win[j] = 1.0f;
            }
        }
    }
}

template <int F, int MAXD, typename TT = float>
__device__ __forceinline__ half8 encode_window(const EncProgram& P, const TT* __restrict__ params,
                                               const float* __restrict__ q, uint32_t w0) {
    float win[8];
    encode_window_f32<F, MAXD, TT>(P, params, q, w0, win);
    half8 h;
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = (half_t)win[j];
    return h;
}

__device__ __forceinline__ floatx4 mfma16(const half8& a, const half8& b, const floatx4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Activated accumulator tiles 2c and 2c+1 as the B fragment of chunk c of the next layer.
__device__ __forceinline__ half8 pack_act(const floatx4& lo, const floatx4& hi, int act) {
    half8 r;
    if (act == ACT_RELU) {
        // ReLU on the rounded halves, two per instruction: rounding keeps the
        // sign and max(NaN, 0) = 0, so only the sign of a zero can differ from
        // rounding max(v, 0)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            r[i] = (half_t)lo[i];
            r[i + 4] = (half_t)hi[i];
        }
        return __builtin_elementwise_max(r, half8{0, 0, 0, 0, 0, 0, 0, 0});
    }
    float v[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    act_fwd_n<8>(act, v);
#pragma unroll
    for (int i = 0; i < 8; ++i) r[i] = (half_t)v[i];
    return r;
}

// Hidden layers + output layer of one 16-query tile, from the first layer's
// accumulators.  Returns the output accumulator (outputs 4g..4g+3 of query lane%16).
template <int W>
__device__ __forceinline__ floatx4 mlp_tail16(const half_t* __restrict__ frag, floatx4* acc, uint32_t L, int act) {
    constexpr int MT = W / 16, KC = W / 32;
    const int lane = threadIdx.x & 63;
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


// Calls f(W, F, MAXD) with integral constants matching the layout.
template <typename Fn>
void dispatch_fused(const Layout& L, Fn&& f) {
    uint32_t F = 2, maxdeg = 0;
    for (const Segment& s : L.segs) {
        if (s.type == SEG_GRID) F = s.F;
        if (s.type == SEG_SH && s.degree > maxdeg) maxdeg = s.degree;
    }
    auto with_d = [&](auto Wc, auto Fc) {
        if (maxdeg <= 4) f(Wc, Fc, std::integral_constant<int, 4>{});
        else f(Wc, Fc, std::integral_constant<int, 8>{});
    };
    auto with_f = [&](auto Wc) {
        switch (F) {
            case 1: with_d(Wc, std::integral_constant<int, 1>{}); break;
            case 2: with_d(Wc, std::integral_constant<int, 2>{}); break;
            case 4: with_d(Wc, std::integral_constant<int, 4>{}); break;
            case 8: with_d(Wc, std::integral_constant<int, 8>{}); break;
            default: fail(NLOD_ERR_UNSUPPORTED, "unsupported n_features_per_level");
        }
    };
    switch (L.net.width) {
        case 32: with_f(std::integral_constant<int, 32>{}); break;
        case 64: with_f(std::integral_constant<int, 64>{}); break;
        case 128: with_f(std::integral_constant<int, 128>{}); break;
        default: fail(NLOD_ERR_UNSUPPORTED, "FP16 MFMA kernels support widths 32, 64 and 128");
    }
}

}  // namespace nlod
