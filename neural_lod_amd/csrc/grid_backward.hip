// Hash-grid parameter gradients without global atomics (gfx950).
//
// The reference scatters w * dL/dy of every (query, level, corner) into the
// table with atomicAdd (tiny-cuda-nn encodings/grid.h:394-517, invoked from
// backward_impl :1102-1193).  On MI355X random FP32 global atomics run at about
// 21 G/s whatever their scope or XCD placement, and FP32 LDS atomics at about
// 0.2 T/s, while 64-bit integer LDS atomics reach about 2.4 T/s.  So the same
// contributions are binned by table tile and summed in LDS in fixed point:
//
//   1. count    per (query block, table tile) how many records fall in the
//               tile; also max |dL/dy| of the pass (sets the fixed-point scale)
//   2. scan     tile-major exclusive scan -> every block's slot range per tile
//   3. scatter  records (entry within the tile; F contributions) to their slots,
//               each record one vector of F + 1 words for F = 2 (12 bytes,
//               unpadded: the reduce streams them) and of next_pow2(F + 1)
//               words otherwise, so a record is one store and one load
//   4. reduce   one workgroup per group of <= kGroupRecs records of one tile:
//               contributions as 64-bit fixed point (2^S units, S from the
//               max so that no sum can overflow) added with ds_add_u64; a tile
//               with one group writes the gradient directly, others a partial
//   5. apply    tiles with several groups (exact integer sum of the partials)
//               and, on overwrite, tiles without contributions
//
// Integer sums make the result independent of the order of the additions:
// the gradient is deterministic.  Contributions below 2^-36 of the pass's
// largest |dL/dy| round to zero (FP32 itself resolves 2^-24 of a value).
// Tiles hold kTileFloats / F entries of one level of one table; segments that
// share a table (RepeatedComposite) share its tiles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "device_common.h"
#include "grid_backward.h"

namespace nlod {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kQueriesPerThread = 8;  // count / scatter: loads of 8 queries in flight per thread
constexpr uint32_t kQueriesPerBlock = kBlock * kQueriesPerThread;
constexpr uint32_t kTileFloats = 8192;      // values of one tile (kTileFloats / F entries): 64 KiB as int64
constexpr uint32_t kGroupRecs = 65536;      // records summed by one reduce workgroup
constexpr uint32_t kInFlight = 8;           // records loaded per thread before their LDS adds
constexpr int kFixedHeadroom = 26;          // log2 of the records one entry can receive in a pass, bounded
constexpr uint32_t kMaxTilesPerRow = 2048;  // LDS counters per (segment, level)

}  // namespace

// One (grid segment, level): what the count / scatter passes need.
struct GridBwdRow {
    float scale;
    uint32_t res, hs;        // resolution, entries of the level
    uint32_t in_offset;      // first input column of the 3 positions
    uint32_t out_row;        // d_enc row of feature 0 of this level
    uint32_t bin_base;       // first tile of this (table, level)
    uint32_t n_tiles;
    int32_t grid_type, interp;
    uint32_t corners;        // 8 (linear / smoothstep) or 1 (nearest)
    uint32_t count_base;     // count row of (tile 0, repetition 0) of this level
    uint32_t reps, rep;      // repetitions reading this table, and which one this row is
};

// One tile: where its entries live in the parameter vector.
struct GridBwdBin {
    uint64_t float_offset;  // parameter index of the tile's first float
    uint32_t n_entries;     // valid entries (the last tile of a level may be short)
    uint32_t count_row;     // first of its `reps` count rows (one per repetition)
    uint32_t reps;
    uint32_t pad_;
};

namespace {

// d_enc rows have stride ld (the full batch), q indexes within the pass
template <int F>
__device__ __forceinline__ bool load_g(const float* __restrict__ d_enc, uint32_t row, size_t ld, uint32_t q,
                                       float* g) {
    bool any = false;
#pragma unroll
    for (int f = 0; f < F; ++f) {
        g[f] = d_enc[(size_t)(row + f) * ld + q];
        any |= g[f] != 0.0f;
    }
    return any;
}

// Calls emit(idx, weight) for every corner the reference interpolates (grid.h:274-341).
template <typename Emit>
__device__ __forceinline__ void corners(const GridBwdRow& r, const float* x, Emit&& emit) {
    float fr[3];
    uint32_t cell[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) grid_pos(x[d], r.scale, r.interp, fr[d], cell[d]);
    if (r.interp == INTERP_NEAREST) {
        emit(grid_index(r.grid_type, r.hs, r.res, cell[0], cell[1], cell[2]), 1.0f);
        return;
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float w = 1.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) w *= ((c >> d) & 1) ? fr[d] : 1.0f - fr[d];
        emit(grid_index(r.grid_type, r.hs, r.res, cell[0] + (c & 1), cell[1] + ((c >> 1) & 1), cell[2] + ((c >> 2) & 1)),
             w);
    }
}

// Pass 1: counts[bin][qblock] and meta[1] = max |dL/dy| bits.  grid (query blocks, rows).
template <int F>
__global__ __launch_bounds__(kBlock) void count_kernel(uint32_t N, uint32_t ld, uint32_t n_qblocks, uint32_t in_stride,
                                                       const GridBwdRow* __restrict__ rows,
                                                       const float* __restrict__ in, const float* __restrict__ d_enc,
                                                       uint32_t* __restrict__ counts, uint32_t* __restrict__ meta) {
    constexpr uint32_t E = kTileFloats / F;
    __shared__ uint32_t cnt[kMaxTilesPerRow];
    __shared__ uint32_t wmax[kBlock / 64];
    const GridBwdRow r = rows[blockIdx.y];
    for (uint32_t t = threadIdx.x; t < r.n_tiles; t += kBlock) cnt[t] = 0;
    __syncthreads();
    // |g| as bits: for non-negative floats the integer order is the float order
    // (NaN sorts above infinity, so a non-finite value wins the max)
    uint32_t mx = 0;
    const uint32_t q0 = blockIdx.x * kQueriesPerBlock;
    // every query's gradient and position loads first, then the index work
    float g[kQueriesPerThread][F], x[kQueriesPerThread][3];
    bool live[kQueriesPerThread];
#pragma unroll
    for (uint32_t k = 0; k < kQueriesPerThread; ++k) {
        const uint32_t q = q0 + k * kBlock + threadIdx.x;
        live[k] = q < N;
        const uint32_t qc = live[k] ? q : N - 1;
        live[k] = load_g<F>(d_enc, r.out_row, ld, qc, g[k]) && live[k];
#pragma unroll
        for (int d = 0; d < 3; ++d) x[k][d] = in[(size_t)qc * in_stride + r.in_offset + d];
    }
#pragma unroll
    for (uint32_t k = 0; k < kQueriesPerThread; ++k) {
        if (!live[k]) continue;
#pragma unroll
        for (int f = 0; f < F; ++f) mx = max(mx, __float_as_uint(fabsf(g[k][f])));
        corners(r, x[k], [&](uint32_t idx, float) { atomicAdd(&cnt[idx / E], 1u); });
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_down((int)mx, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = 0;
        for (int w = 0; w < kBlock / 64; ++w) m = max(m, wmax[w]);
        if (m) atomicMax(&meta[1], m);
    }
    for (uint32_t t = threadIdx.x; t < r.n_tiles; t += kBlock)
        counts[(size_t)(r.count_base + t * r.reps + r.rep) * n_qblocks + blockIdx.x] = cnt[t];
}

__global__ void reset_meta_kernel(uint32_t* meta) {
    if (threadIdx.x < 4) meta[threadIdx.x] = 0;
}

// Pass 2a: per tile, exclusive scan over (repetition, query block) in place,
// and the tile total.  grid: tiles.
__global__ __launch_bounds__(kBlock) void scan_rows_kernel(uint32_t n_qblocks, const GridBwdBin* __restrict__ bins,
                                                           uint32_t* __restrict__ counts,
                                                           uint32_t* __restrict__ totals) {
    __shared__ uint32_t part[kBlock];
    const GridBwdBin bin = bins[blockIdx.x];
    uint32_t* row = counts + (size_t)bin.count_row * n_qblocks;
    const uint32_t len = bin.reps * n_qblocks;
    // each thread owns a contiguous run of the row
    const uint32_t per = (len + kBlock - 1) / kBlock;
    const uint32_t b0 = min(len, threadIdx.x * per), b1 = min(len, b0 + per);
    uint32_t s = 0;
    for (uint32_t i = b0; i < b1; ++i) s += row[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < kBlock; ++i) {
            const uint32_t v = part[i];
            part[i] = acc;
            acc += v;
        }
        totals[blockIdx.x] = acc;
    }
    __syncthreads();
    uint32_t acc = part[threadIdx.x];
    for (uint32_t i = b0; i < b1; ++i) {
        const uint32_t v = row[i];
        row[i] = acc;
        acc += v;
    }
}

// Pass 2b (one block): tile starts and group numbering (a group = up to
// chunk_recs records of one tile).  meta[0] = groups.
__device__ __forceinline__ uint32_t groups_of(uint32_t total, uint32_t chunk_recs) {
    return (total + chunk_recs - 1) / chunk_recs;
}

__global__ __launch_bounds__(kBlock) void scan_bins_kernel(uint32_t n_bins, uint32_t chunk_recs,
                                                           const uint32_t* __restrict__ totals,
                                                           uint32_t* __restrict__ bin_start,
                                                           uint32_t* __restrict__ group_start,
                                                           uint32_t* __restrict__ meta) {
    __shared__ uint32_t part_r[kBlock], part_c[kBlock];
    const uint32_t per = (n_bins + kBlock - 1) / kBlock;
    const uint32_t b0 = min(n_bins, threadIdx.x * per), b1 = min(n_bins, b0 + per);
    uint32_t sr = 0, sc = 0;
    for (uint32_t i = b0; i < b1; ++i) {
        sr += totals[i];
        sc += groups_of(totals[i], chunk_recs);
    }
    part_r[threadIdx.x] = sr;
    part_c[threadIdx.x] = sc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ar = 0, ac = 0;
        for (int i = 0; i < kBlock; ++i) {
            const uint32_t vr = part_r[i], vc = part_c[i];
            part_r[i] = ar;
            part_c[i] = ac;
            ar += vr;
            ac += vc;
        }
        meta[0] = ac;
        group_start[n_bins] = ac;
    }
    __syncthreads();
    uint32_t ar = part_r[threadIdx.x], ac = part_c[threadIdx.x];
    for (uint32_t i = b0; i < b1; ++i) {
        bin_start[i] = ar;
        group_start[i] = ac;
        ar += totals[i];
        ac += groups_of(totals[i], chunk_recs);
    }
}

// Words per record: the key and F values; padded to a power of two except for
// F = 2, whose 3-word records load and store as one dwordx3.
template <int F>
struct RecWords {
    static constexpr int value = F == 2 ? 3 : F + 1 <= 2 ? 2 : F + 1 <= 4 ? 4 : F + 1 <= 8 ? 8 : 16;
};

template <int RW>
__device__ __forceinline__ void store_rec(uint32_t* __restrict__ dst, const uint32_t (&w)[RW]) {
    if constexpr (RW == 2) {
        *reinterpret_cast<uint2*>(dst) = make_uint2(w[0], w[1]);
    } else if constexpr (RW == 3) {
        *reinterpret_cast<uint3*>(dst) = make_uint3(w[0], w[1], w[2]);
    } else {
#pragma unroll
        for (int i = 0; i < RW; i += 4)
            *reinterpret_cast<uint4*>(dst + i) = make_uint4(w[i], w[i + 1], w[i + 2], w[i + 3]);
    }
}

template <int RW>
__device__ __forceinline__ void load_rec(const uint32_t* __restrict__ src, uint32_t (&w)[RW]) {
    if constexpr (RW == 2) {
        const uint2 a = *reinterpret_cast<const uint2*>(src);
        w[0] = a.x;
        w[1] = a.y;
    } else if constexpr (RW == 3) {
        const uint3 a = *reinterpret_cast<const uint3*>(src);
        w[0] = a.x;
        w[1] = a.y;
        w[2] = a.z;
    } else {
#pragma unroll
        for (int i = 0; i < RW; i += 4) {
            const uint4 a = *reinterpret_cast<const uint4*>(src + i);
            w[i] = a.x;
            w[i + 1] = a.y;
            w[i + 2] = a.z;
            w[i + 3] = a.w;
        }
    }
}

// Pass 3: record of slot i = words [i * RW, (i + 1) * RW): entry within the
// tile, then the F contributions.
template <int F>
__global__ __launch_bounds__(kBlock) void scatter_kernel(uint32_t N, uint32_t ld, uint32_t n_qblocks, uint32_t in_stride,
                                                         const GridBwdRow* __restrict__ rows,
                                                         const float* __restrict__ in,
                                                         const float* __restrict__ d_enc,
                                                         const uint32_t* __restrict__ row_offsets,
                                                         const uint32_t* __restrict__ bin_start,
                                                         uint32_t* __restrict__ recs) {
    constexpr uint32_t E = kTileFloats / F;
    constexpr int RW = RecWords<F>::value;
    __shared__ uint32_t cursor[kMaxTilesPerRow];
    const GridBwdRow r = rows[blockIdx.y];
    for (uint32_t t = threadIdx.x; t < r.n_tiles; t += kBlock) {
        const uint32_t b = r.bin_base + t;
        cursor[t] = bin_start[b] + row_offsets[(size_t)(r.count_base + t * r.reps + r.rep) * n_qblocks + blockIdx.x];
    }
    __syncthreads();
    const uint32_t q0 = blockIdx.x * kQueriesPerBlock;
    // every query's gradient and position loads first, then the index work
    float gq[kQueriesPerThread][F], xq[kQueriesPerThread][3];
    bool live[kQueriesPerThread];
#pragma unroll
    for (uint32_t k = 0; k < kQueriesPerThread; ++k) {
        const uint32_t q = q0 + k * kBlock + threadIdx.x;
        live[k] = q < N;
        const uint32_t qc = live[k] ? q : N - 1;
        live[k] = load_g<F>(d_enc, r.out_row, ld, qc, gq[k]) && live[k];
#pragma unroll
        for (int d = 0; d < 3; ++d) xq[k][d] = in[(size_t)qc * in_stride + r.in_offset + d];
    }
#pragma unroll
    for (uint32_t k = 0; k < kQueriesPerThread; ++k) {
        if (!live[k]) continue;
        const float* g = gq[k];
        corners(r, xq[k], [&](uint32_t idx, float w) {
            const uint32_t slot = atomicAdd(&cursor[idx / E], 1u);
            uint32_t rec[RW];
            rec[0] = idx % E;
#pragma unroll
            for (int f = 0; f < F; ++f) rec[1 + f] = __float_as_uint(w * g[f]);
#pragma unroll
            for (int i = F + 1; i < RW; ++i) rec[i] = 0u;
            store_rec<RW>(recs + (size_t)slot * RW, rec);
        });
    }
}

// Fixed-point exponent of a pass: contributions are scaled by 2^S with
// max * 2^S < 2^(63 - kFixedHeadroom), so sums of up to 2^kFixedHeadroom of them fit.
__device__ __forceinline__ int fixed_shift(uint32_t max_bits) {
    int e = 0;
    (void)frexpf(__uint_as_float(max_bits), &e);  // max < 2^e
    return 62 - kFixedHeadroom - e;
}

__device__ __forceinline__ bool finite_bits(uint32_t b) { return b < 0x7f800000u; }

// Pass 4: group gi of its tile.  grid: upper bound of groups.
template <int F>
__global__ __launch_bounds__(kBlock) void reduce_kernel(uint32_t n_bins, const uint32_t* __restrict__ meta,
                                                        const uint32_t* __restrict__ totals,
                                                        const uint32_t* __restrict__ bin_start,
                                                        const uint32_t* __restrict__ group_start,
                                                        const GridBwdBin* __restrict__ bins,
                                                        const uint32_t* __restrict__ recs,
                                                        long long* __restrict__ partial, float* __restrict__ grads,
                                                        int overwrite) {
    constexpr int RW = RecWords<F>::value;
    __shared__ unsigned long long acc[kTileFloats];
    const uint32_t gi = blockIdx.x;
    if (gi >= meta[0]) return;
    // tile of group gi: the last bin whose first group is <= gi
    uint32_t lo = 0, hi = n_bins;  // invariant: group_start[lo] <= gi < group_start[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (group_start[mid] <= gi) lo = mid;
        else hi = mid;
    }
    const uint32_t b = lo;
    const uint32_t begin = bin_start[b] + (gi - group_start[b]) * kGroupRecs;
    const uint32_t end = min(bin_start[b] + totals[b], begin + kGroupRecs);
    const uint32_t mbits = meta[1];
    const bool finite = finite_bits(mbits);
    const int S = finite ? fixed_shift(mbits) : 0;
    for (uint32_t i = threadIdx.x; i < kTileFloats; i += kBlock) acc[i] = 0ull;
    __syncthreads();
    // software-pipelined: the next batch's loads are issued before this
    // batch's LDS adds (lanes past the end load the last record and add 0)
    uint32_t cur[kInFlight][RW] = {}, nxt[kInFlight][RW] = {};  // nxt stays unread after the last batch
    auto load_batch = [&](uint32_t i0, uint32_t (&dst)[kInFlight][RW]) {
#pragma unroll
        for (uint32_t k = 0; k < kInFlight; ++k) {
            const uint32_t i = min(i0 + k * kBlock + threadIdx.x, end - 1);
            load_rec<RW>(recs + (size_t)i * RW, dst[k]);
        }
    };
    if (begin < end) load_batch(begin, cur);
    for (uint32_t i0 = begin; i0 < end; i0 += kBlock * kInFlight) {
        const uint32_t i1 = i0 + kBlock * kInFlight;
        if (i1 < end) load_batch(i1, nxt);
#pragma unroll
        for (uint32_t k = 0; k < kInFlight; ++k) {
            const bool valid = i0 + k * kBlock + threadIdx.x < end;
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const long long q = valid ? __float2ll_rn(ldexpf(__uint_as_float(cur[k][1 + f]), S)) : 0ll;
                atomicAdd(&acc[cur[k][0] * F + f], (unsigned long long)q);
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kInFlight; ++k)
#pragma unroll
            for (int j = 0; j < RW; ++j) cur[k][j] = nxt[k][j];
    }
    __syncthreads();
    const GridBwdBin bin = bins[b];
    const uint32_t n_vals = bin.n_entries * F;
    if (group_start[b + 1] - group_start[b] == 1) {
        // the tile's only group: this is the gradient
        for (uint32_t i = threadIdx.x; i < n_vals; i += kBlock) {
            const float g = finite ? ldexpf((float)(long long)acc[i], -S) : __uint_as_float(0x7fc00000u);
            float* dst = grads + bin.float_offset + i;
            if (overwrite) *dst = g;
            else *dst += g;
        }
    } else {
        long long* out = partial + (size_t)gi * kTileFloats;
        for (uint32_t i = threadIdx.x; i < n_vals; i += kBlock) out[i] = (long long)acc[i];
    }
}

// Pass 5: tiles with several groups, and (overwrite) tiles with none.  grid (tiles, kTileFloats / kBlock).
__global__ __launch_bounds__(kBlock) void apply_kernel(const GridBwdBin* __restrict__ bins,
                                                       const uint32_t* __restrict__ group_start,
                                                       const uint32_t* __restrict__ meta,
                                                       const long long* __restrict__ partial,
                                                       float* __restrict__ grads, uint32_t F, int overwrite) {
    const uint32_t b = blockIdx.x;
    const uint32_t c0 = group_start[b], c1 = group_start[b + 1];
    if (c1 - c0 == 1 || (c1 == c0 && !overwrite)) return;
    const uint32_t e = blockIdx.y * kBlock + threadIdx.x;
    const GridBwdBin bin = bins[b];
    if (e >= bin.n_entries * F) return;
    float* dst = grads + bin.float_offset + e;
    if (c1 == c0) {
        *dst = 0.0f;
        return;
    }
    long long s = 0;
    for (uint32_t c = c0; c < c1; ++c) s += partial[(size_t)c * kTileFloats + e];
    const uint32_t mbits = meta[1];
    const float g = finite_bits(mbits) ? ldexpf((float)s, -fixed_shift(mbits)) : __uint_as_float(0x7fc00000u);
    if (overwrite) *dst = g;
    else *dst += g;
}

template <typename Fn>
void by_F(uint32_t F, Fn&& f) {
    switch (F) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        default: fail(NLOD_ERR_UNSUPPORTED, "unsupported n_features_per_level");
    }
}

}  // namespace

struct GridBackward::HostPlan {
    std::vector<GridBwdRow> rows;
    std::vector<GridBwdBin> bins;
};

GridBackward::GridBackward(const Layout& L) : n_input_dims_(L.n_input_dims) {
    // tables are identified by their parameter offset (shared by repetitions)
    struct Table {
        uint64_t param_offset;
        uint32_t reps;
        std::vector<uint32_t> level_bin_base;
    };
    std::vector<Table> tables;
    for (const Segment& s : L.segs) {
        if (s.type != SEG_GRID) continue;
        bool found = false;
        for (Table& tb : tables)
            if (tb.param_offset == s.param_offset) {
                ++tb.reps;
                found = true;
            }
        if (!found) tables.push_back(Table{s.param_offset, 1, {}});
    }
    auto plan = std::make_unique<HostPlan>();
    std::vector<uint32_t> rep_seen(tables.size(), 0);
    uint32_t F = 0;
    for (const Segment& s : L.segs) {
        if (s.type != SEG_GRID) continue;
        if (F && s.F != F) return;  // mixed feature counts: not supported here
        F = s.F;
        if (F != 1 && F != 2 && F != 4 && F != 8) return;
        const uint32_t E = kTileFloats / F;
        size_t ti = 0;
        while (tables[ti].param_offset != s.param_offset) ++ti;
        Table& t = tables[ti];
        if (t.level_bin_base.empty()) {
            for (uint32_t l = 0; l < s.n_levels; ++l) {
                const uint32_t hs = s.levels.offset[l + 1] - s.levels.offset[l];
                const uint32_t n_tiles = (hs + E - 1) / E;
                if (n_tiles > kMaxTilesPerRow) return;
                t.level_bin_base.push_back((uint32_t)plan->bins.size());
                for (uint32_t k = 0; k < n_tiles; ++k) {
                    GridBwdBin b{};
                    b.float_offset = s.param_offset + ((uint64_t)s.levels.offset[l] + (uint64_t)k * E) * F;
                    b.n_entries = std::min(E, hs - k * E);
                    b.count_row = count_rows_;
                    b.reps = t.reps;
                    count_rows_ += t.reps;
                    plan->bins.push_back(b);
                }
            }
        }
        const uint32_t rep = rep_seen[ti]++;
        for (uint32_t l = 0; l < s.n_levels; ++l) {
            GridBwdRow r{};
            r.scale = s.levels.scale[l];
            r.res = s.levels.res[l];
            r.hs = s.levels.offset[l + 1] - s.levels.offset[l];
            r.in_offset = s.in_offset;
            r.out_row = s.out_offset + l * F;
            r.bin_base = t.level_bin_base[l];
            r.n_tiles = (r.hs + E - 1) / E;
            r.grid_type = s.grid_type;
            r.interp = s.interp;
            r.corners = s.interp == INTERP_NEAREST ? 1u : 8u;
            r.count_base = plan->bins[r.bin_base].count_row;
            r.reps = t.reps;
            r.rep = rep;
            plan->rows.push_back(r);
        }
    }
    if (plan->rows.empty()) return;
    F_ = F;
    n_rows_ = (uint32_t)plan->rows.size();
    n_bins_ = (uint32_t)plan->bins.size();
    for (const GridBwdRow& r : plan->rows) corners_per_query_ += r.corners;
    host_ = plan.release();
    supported_ = true;
}

GridBackward::~GridBackward() { delete host_; }

uint64_t GridBackward::max_groups(uint32_t n) const {
    // sum over tiles of ceil(t / G) <= R / G + tiles
    return (uint64_t)n * corners_per_query_ / kGroupRecs + n_bins_;
}

uint32_t GridBackward::queries_per_pass() const {
    // at most 2^27 records per pass (1.5 GiB at F = 2); an entry then receives at
    // most 2^27 / (records per query and level) <= 2^kFixedHeadroom contributions
    const uint64_t cap = (1ull << 27) / std::max<uint32_t>(corners_per_query_, 1);
    return (uint32_t)std::max<uint64_t>(kQueriesPerBlock, cap / kQueriesPerBlock * kQueriesPerBlock);
}

void GridBackward::prepare(uint32_t N) {
    const uint32_t n = std::min(N, queries_per_pass());
    const uint32_t nqb = div_round_up(n, kQueriesPerBlock);
    const uint64_t max_records = (uint64_t)n * corners_per_query_;
    counts_.ensure(sizeof(uint32_t) * (size_t)count_rows_ * nqb);
    totals_.ensure(sizeof(uint32_t) * n_bins_);
    bin_start_.ensure(sizeof(uint32_t) * n_bins_);
    chunk_start_.ensure(sizeof(uint32_t) * (n_bins_ + 1));
    meta_.ensure(sizeof(uint32_t) * 4);
    const size_t rec_words = F_ == 2 ? 3 : F_ + 1 <= 2 ? 2 : F_ + 1 <= 4 ? 4 : F_ + 1 <= 8 ? 8 : 16;
    records_.ensure(sizeof(uint32_t) * (size_t)max_records * rec_words);
    partial_.ensure(sizeof(long long) * (size_t)max_groups(n) * kTileFloats);
}

void GridBackward::run(hipStream_t s, const float* in, uint32_t N, const float* d_enc, float* grads,
                       bool overwrite) {
    if (!supported_) fail(NLOD_ERR_STATE, "binned grid backward not available for this layout");
    if (N == 0) return;
    if (!uploaded_) {
        d_rows_.ensure(sizeof(GridBwdRow) * host_->rows.size());
        d_bins_.ensure(sizeof(GridBwdBin) * host_->bins.size());
        NLOD_HIP(hipMemcpy(d_rows_.p, host_->rows.data(), sizeof(GridBwdRow) * host_->rows.size(),
                           hipMemcpyHostToDevice));
        NLOD_HIP(hipMemcpy(d_bins_.p, host_->bins.data(), sizeof(GridBwdBin) * host_->bins.size(),
                           hipMemcpyHostToDevice));
        uploaded_ = true;
    }
    const uint32_t pass = queries_per_pass();
    prepare(N);
    for (uint32_t q0 = 0; q0 < N; q0 += pass) {
        const uint32_t n = std::min(pass, N - q0);
        const uint32_t nqb = div_round_up(n, kQueriesPerBlock);
        const uint64_t n_groups = max_groups(n);
        const float* in_c = in + (size_t)q0 * n_input_dims_;
        // d_enc is SoA [K][N]: the pass's columns start at q0 within each row of stride N
        const float* denc_c = d_enc + q0;
        hipLaunchKernelGGL(reset_meta_kernel, dim3(1), dim3(64), 0, s, meta_.as<uint32_t>());
        NLOD_LAUNCH_CHECK();
        by_F(F_, [&](auto Fc) {
            constexpr int F = decltype(Fc)::value;
            hipLaunchKernelGGL(count_kernel<F>, dim3(nqb, n_rows_), dim3(kBlock), 0, s, n, N, nqb, n_input_dims_,
                               d_rows_.as<GridBwdRow>(), in_c, denc_c, counts_.as<uint32_t>(), meta_.as<uint32_t>());
        });
        NLOD_LAUNCH_CHECK();
        hipLaunchKernelGGL(scan_rows_kernel, dim3(n_bins_), dim3(kBlock), 0, s, nqb, d_bins_.as<GridBwdBin>(),
                           counts_.as<uint32_t>(), totals_.as<uint32_t>());
        NLOD_LAUNCH_CHECK();
        hipLaunchKernelGGL(scan_bins_kernel, dim3(1), dim3(kBlock), 0, s, n_bins_, kGroupRecs,
                           totals_.as<uint32_t>(), bin_start_.as<uint32_t>(), chunk_start_.as<uint32_t>(),
                           meta_.as<uint32_t>());
        NLOD_LAUNCH_CHECK();
        const int ow = (overwrite && q0 == 0) ? 1 : 0;
        by_F(F_, [&](auto Fc) {
            constexpr int F = decltype(Fc)::value;
            hipLaunchKernelGGL(scatter_kernel<F>, dim3(nqb, n_rows_), dim3(kBlock), 0, s, n, N, nqb, n_input_dims_,
                               d_rows_.as<GridBwdRow>(), in_c, denc_c, counts_.as<uint32_t>(),
                               bin_start_.as<uint32_t>(), records_.as<uint32_t>());
            NLOD_LAUNCH_CHECK();
            hipLaunchKernelGGL(reduce_kernel<F>, dim3((uint32_t)n_groups), dim3(kBlock), 0, s, n_bins_,
                               meta_.as<uint32_t>(), totals_.as<uint32_t>(), bin_start_.as<uint32_t>(),
                               chunk_start_.as<uint32_t>(), d_bins_.as<GridBwdBin>(), records_.as<uint32_t>(),
                               partial_.as<long long>(), grads, ow);
        });
        NLOD_LAUNCH_CHECK();
        hipLaunchKernelGGL(apply_kernel, dim3(n_bins_, kTileFloats / kBlock), dim3(kBlock), 0, s,
                           d_bins_.as<GridBwdBin>(), chunk_start_.as<uint32_t>(), meta_.as<uint32_t>(),
                           partial_.as<long long>(), grads, F_, ow);
        NLOD_LAUNCH_CHECK();
    }
}

}  // namespace nlod
