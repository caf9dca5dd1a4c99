// Hash-grid parameter gradients without global atomics (gfx950).
//
// The reference scatters w * dL/dy of every (query, level, corner) into the
// table with atomicAdd (tiny-cuda-nn encodings/grid.h:394-517, invoked from
// backward_impl :1102-1193).  On MI355X random FP32 global atomics run at about
// 21 G/s whatever their scope or XCD placement, and FP32 LDS atomics at about
// 0.2 T/s, while 64-bit integer LDS atomics reach about 2.4 T/s.  So the same
// contributions are binned by table tile and summed in LDS in fixed point:
//
//   1. bin      per (query block, level): every x-edge of every query's cell
//               (the corners x0 and x0 + 1 of one (y, z) row) becomes one
//               record in its table tile -- or two single-corner records when
//               the corners lie in different tiles.  The block counts its
//               records per tile in LDS, gives each tile a run of its own
//               region of the record buffer, and writes the records there;
//               per (tile, block) it stores the run's length and address.  It
//               also raises max |dL/dy| of the pass (sets the fixed-point scale).
//   2. scan     per tile, exclusive scan of its run lengths over (repetition,
//               query block) -> the tile's records in one order, cut into
//               groups of <= kGroupRecs
//   3. reduce   one workgroup per group: each record's contributions as
//               64-bit fixed point (2^S units, S from the max so that no sum
//               can overflow) added with ds_add_u64; a tile with one group
//               writes the gradient directly, others a partial
//   4. apply    tiles with several groups (exact integer sum of the partials)
//               and, on overwrite, tiles without contributions
//
// A record is 4 words for F <= 2: key, weight, then F values a_f.  An edge
// record (key bit 31 set, entries e0 and e1 of the tile in bits 0-12 and
// 13-25) contributes (1 - fx) a_f to e0 and fx a_f to e1, with a_f = (w_y w_z)
// g_f and fx the x fraction of the cell; a single record (bit 31 clear, entry
// e0) contributes weight * a_f.  An edge record moves 16 bytes for two
// corners where one record per corner moved 12 per corner.
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
constexpr uint32_t kMaxRuns = 1024;         // runs of one tile: repetitions x query blocks of a pass
constexpr uint32_t kEdgeFlag = 1u << 31;

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

// The cell of query x at row r as x-edges (grid.h:274-341): for each of the
// 4 (y, z) rows of the cell, emit(i0, i1, fx, wyz) with the entries of the
// corners x0 and x0 + 1, the x fraction and the weight w_y w_z; nearest
// interpolation gives one edge whose two corners coincide (fx = 0, weight 1).
template <typename Emit>
__device__ __forceinline__ void edges(const GridBwdRow& r, const float* x, Emit&& emit) {
    float fr[3];
    uint32_t cell[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) grid_pos(x[d], r.scale, r.interp, fr[d], cell[d]);
    if (r.interp == INTERP_NEAREST) {
        const uint32_t i = grid_index(r.grid_type, r.hs, r.res, cell[0], cell[1], cell[2]);
        emit(i, i, 0.0f, 1.0f);
        return;
    }
    const LevelIndexer li = level_indexer(r.grid_type, r.hs, r.res);
    const uint32_t y0 = cell[1] * li.my, z0 = cell[2] * li.mz;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const uint32_t dy = p & 1, dz = p >> 1;
        const float wyz = (dy ? fr[1] : 1.0f - fr[1]) * (dz ? fr[2] : 1.0f - fr[2]);
        const uint32_t ym = dy ? y0 + li.my : y0, zm = dz ? z0 + li.mz : z0;
        emit(level_index(li, cell[0], ym, zm), level_index(li, cell[0] + 1, ym, zm), fr[0], wyz);
    }
}

// Exclusive prefix sum of v over the block (wave shuffles, then the wave
// totals through s_wave); returns the block's total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t& excl) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        const uint32_t t = s_wave[w];
        if (w < (int)wave) before += t;
        total += t;
    }
    excl = before + x - v;
    return total;
}

// meta: [0] groups, [1] max |dL/dy| bits being raised by the bin pass, [2]
// tiles listed for the apply pass, [3] the pass's max, moved there from [1]
// by the group scan, which leaves [1] zero for the next pass.  Cleared once,
// when the buffer is allocated.
__global__ void reset_meta_kernel(uint32_t* meta) {
    if (threadIdx.x < 4) meta[threadIdx.x] = 0;
}

// Pass 2a: per tile, exclusive scan of its run lengths over (repetition,
// query block) in place, and the tile total.  grid: tiles.
__global__ __launch_bounds__(kBlock) void scan_rows_kernel(uint32_t n_qblocks, const GridBwdBin* __restrict__ bins,
                                                           uint32_t* __restrict__ counts,
                                                           uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_wave[kBlock / 64];
    const GridBwdBin bin = bins[blockIdx.x];
    uint32_t* row = counts + (size_t)bin.count_row * n_qblocks;
    const uint32_t len = bin.reps * n_qblocks;
    // each thread owns a contiguous run of the row
    const uint32_t per = (len + kBlock - 1) / kBlock;
    const uint32_t b0 = min(len, threadIdx.x * per), b1 = min(len, b0 + per);
    uint32_t s = 0;
    for (uint32_t i = b0; i < b1; ++i) s += row[i];
    uint32_t acc;
    const uint32_t total = block_exclusive_scan(s, s_wave, acc);
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
    for (uint32_t i = b0; i < b1; ++i) {
        const uint32_t v = row[i];
        row[i] = acc;
        acc += v;
    }
}

// Pass 2b (one block): group numbering (a group = up to chunk_recs records
// of one tile).  meta[0] = groups.  Also lists, in order, the tiles the apply
// pass has work for (several groups, or none when overwriting): meta[2] of them.
// meta[3] = the max the bin pass raised in meta[1], which becomes 0 again.
__device__ __forceinline__ uint32_t groups_of(uint32_t total, uint32_t chunk_recs) {
    return (total + chunk_recs - 1) / chunk_recs;
}

__device__ __forceinline__ bool needs_apply(uint32_t groups, int overwrite) {
    return groups >= 2 || (groups == 0 && overwrite);
}

__global__ __launch_bounds__(kBlock) void scan_bins_kernel(uint32_t n_bins, uint32_t chunk_recs, int overwrite,
                                                           const uint32_t* __restrict__ totals,
                                                           uint32_t* __restrict__ group_start,
                                                           uint32_t* __restrict__ apply_list,
                                                           uint32_t* __restrict__ meta) {
    // both counts scanned at once: groups in the high word, listed tiles in the
    // low word (their sum stays below 2^32, so no carry crosses)
    __shared__ uint64_t s_wave[kBlock / 64];
    const uint32_t per = (n_bins + kBlock - 1) / kBlock;
    const uint32_t b0 = min(n_bins, threadIdx.x * per), b1 = min(n_bins, b0 + per);
    uint64_t mine = 0;
    for (uint32_t i = b0; i < b1; ++i) {
        const uint32_t g = groups_of(totals[i], chunk_recs);
        mine += ((uint64_t)g << 32) | (needs_apply(g, overwrite) ? 1u : 0u);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t x = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint64_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        const uint64_t v = s_wave[w];
        if (w < wave) before += v;
        total += v;
    }
    const uint64_t excl = before + x - mine;
    if (threadIdx.x == 0) {
        meta[0] = (uint32_t)(total >> 32);
        meta[2] = (uint32_t)total;
        meta[3] = meta[1];
        meta[1] = 0;
        group_start[n_bins] = (uint32_t)(total >> 32);
    }
    uint32_t ac = (uint32_t)(excl >> 32), aa = (uint32_t)excl;
    for (uint32_t i = b0; i < b1; ++i) {
        const uint32_t g = groups_of(totals[i], chunk_recs);
        group_start[i] = ac;
        ac += g;
        if (needs_apply(g, overwrite)) apply_list[aa++] = i;
    }
}

// Words per record: key, weight and F values, padded to a power of two (one
// dwordx4 for F <= 2).
template <int F>
struct RecWords {
    static constexpr int value = F + 2 <= 4 ? 4 : F + 2 <= 8 ? 8 : 16;
};

template <int RW>
__device__ __forceinline__ void store_rec(uint32_t* __restrict__ dst, const uint32_t (&w)[RW]) {
#pragma unroll
    for (int i = 0; i < RW; i += 4)
        *reinterpret_cast<uint4*>(dst + i) = make_uint4(w[i], w[i + 1], w[i + 2], w[i + 3]);
}

template <int RW>
__device__ __forceinline__ void load_rec(const uint32_t* __restrict__ src, uint32_t (&w)[RW]) {
#pragma unroll
    for (int i = 0; i < RW; i += 4) {
        const uint4 a = *reinterpret_cast<const uint4*>(src + i);
        w[i] = a.x;
        w[i + 1] = a.y;
        w[i + 2] = a.z;
        w[i + 3] = a.w;
    }
}

// Records of one (row, query block) live in a region of kQueriesPerBlock x
// r.corners slots (the most the block can write: every edge split).
__device__ __forceinline__ size_t region_of(uint32_t row, uint32_t qblock, uint32_t n_qblocks) {
    return ((size_t)row * n_qblocks + qblock) * kQueriesPerBlock * 8;
}

// Pass 1: records of (query block, row); run_len / run_addr[count row of
// (tile, repetition)][query block] = the run's record count and its first
// slot; raises meta[1] (0 on entry) to max |dL/dy| bits.  grid (query blocks, rows).
template <int F>
__global__ __launch_bounds__(kBlock) void bin_kernel(uint32_t N, uint32_t ld, uint32_t n_qblocks, uint32_t in_stride,
                                                     const GridBwdRow* __restrict__ rows, const float* __restrict__ in,
                                                     const float* __restrict__ d_enc, uint32_t* __restrict__ run_len,
                                                     uint32_t* __restrict__ run_addr, uint32_t* __restrict__ recs,
                                                     uint32_t* __restrict__ meta) {
    constexpr uint32_t E = kTileFloats / F;
    constexpr int RW = RecWords<F>::value;
    __shared__ uint32_t cnt[kMaxTilesPerRow];
    __shared__ uint32_t s_wave[kBlock / 64];
    __shared__ uint32_t wmax[kBlock / 64];
    const GridBwdRow r = rows[blockIdx.y];
    for (uint32_t t = threadIdx.x; t < r.n_tiles; t += kBlock) cnt[t] = 0;
    __syncthreads();
    // every query's gradient and position loads first, then the index work
    float gq[kQueriesPerThread][F], xq[kQueriesPerThread][3];
    bool live[kQueriesPerThread];
    const uint32_t q0 = blockIdx.x * kQueriesPerBlock;
#pragma unroll
    for (uint32_t k = 0; k < kQueriesPerThread; ++k) {
        const uint32_t q = q0 + k * kBlock + threadIdx.x;
        live[k] = q < N;
        const uint32_t qc = live[k] ? q : N - 1;
        live[k] = load_g<F>(d_enc, r.out_row, ld, qc, gq[k]) && live[k];
#pragma unroll
        for (int d = 0; d < 3; ++d) xq[k][d] = in[(size_t)qc * in_stride + r.in_offset + d];
    }
    // count the records per tile, and |g| as bits: for non-negative floats the
    // integer order is the float order (NaN sorts above infinity, so a
    // non-finite value wins the max)
    uint32_t mx = 0;
#pragma unroll
    for (uint32_t k = 0; k < kQueriesPerThread; ++k) {
        if (!live[k]) continue;
#pragma unroll
        for (int f = 0; f < F; ++f) mx = max(mx, __float_as_uint(fabsf(gq[k][f])));
        edges(r, xq[k], [&](uint32_t i0, uint32_t i1, float, float) {
            const uint32_t t0 = i0 / E, t1 = i1 / E;
            atomicAdd(&cnt[t0], 1u);
            if (t1 != t0) atomicAdd(&cnt[t1], 1u);
        });
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
    // exclusive scan of the tile counts (each thread a contiguous run of tiles)
    // -> the first slot of each tile's run in the block's region; cnt becomes
    // the cursor of the writes
    const uint32_t per = (r.n_tiles + kBlock - 1) / kBlock;
    const uint32_t t_lo = min(r.n_tiles, threadIdx.x * per), t_hi = min(r.n_tiles, t_lo + per);
    uint32_t sum = 0;
    for (uint32_t t = t_lo; t < t_hi; ++t) sum += cnt[t];
    uint32_t acc;
    block_exclusive_scan(sum, s_wave, acc);
    const size_t region = region_of(blockIdx.y, blockIdx.x, n_qblocks);
    for (uint32_t t = t_lo; t < t_hi; ++t) {
        const uint32_t c = cnt[t];
        const size_t at = (size_t)(r.count_base + t * r.reps + r.rep) * n_qblocks + blockIdx.x;
        run_len[at] = c;
        run_addr[at] = (uint32_t)(region + acc);  // slots of one pass stay below 2^32
        cnt[t] = acc;
        acc += c;
    }
    __syncthreads();
    // the records
#pragma unroll
    for (uint32_t k = 0; k < kQueriesPerThread; ++k) {
        if (!live[k]) continue;
        const float* g = gq[k];
        edges(r, xq[k], [&](uint32_t i0, uint32_t i1, float fx, float wyz) {
            const uint32_t t0 = i0 / E, t1 = i1 / E;
            uint32_t rec[RW];
#pragma unroll
            for (int f = 0; f < F; ++f) rec[2 + f] = __float_as_uint(wyz * g[f]);
#pragma unroll
            for (int i = F + 2; i < RW; ++i) rec[i] = 0u;
            if (t0 == t1) {
                rec[0] = kEdgeFlag | (i0 % E) | ((i1 % E) << 13);
                rec[1] = __float_as_uint(fx);
                store_rec<RW>(recs + (region + atomicAdd(&cnt[t0], 1u)) * RW, rec);
            } else {
                rec[0] = i0 % E;
                rec[1] = __float_as_uint(1.0f - fx);
                store_rec<RW>(recs + (region + atomicAdd(&cnt[t0], 1u)) * RW, rec);
                rec[0] = i1 % E;
                rec[1] = __float_as_uint(fx);
                store_rec<RW>(recs + (region + atomicAdd(&cnt[t1], 1u)) * RW, rec);
            }
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

// c * 2^S rounded to the nearest integer (ties to even), as two's-complement
// int64: the rounded value r splits exactly into r = hi * 2^32 + lo with hi =
// floor(r / 2^32) in [-2^30, 2^30) (|r| < 2^62) and lo in [0, 2^32), both
// integral floats that convert exactly; the pair is the int64's two words.
__device__ __forceinline__ unsigned long long to_fixed(float c, int S) {
    const float r = __builtin_rintf(ldexpf(c, S));
    const float hi = floorf(r * 0x1p-32f);
    const float lo = fmaf(hi, -0x1p32f, r);
    const uint32_t w0 = (uint32_t)lo;
    const uint32_t w1 = (uint32_t)(int32_t)hi;
    return ((unsigned long long)w1 << 32) | w0;
}

// Pass 3: group gi of its tile.  The tile's records are its runs in
// (repetition, query block) order: run j holds run_pre[j + 1] - run_pre[j]
// records from slot run_addr[j].  Each wave walks a contiguous quarter of the
// group, 64 records (one per lane) at a time; a lane keeps the run its record
// lies in and moves it forward as its record does (a step of 64 records
// crosses about one run), so the slot of a record costs one or two LDS reads.
// grid: upper bound of groups.
template <int F>
__global__ __launch_bounds__(kBlock) void reduce_kernel(uint32_t n_bins, uint32_t n_qblocks,
                                                        const uint32_t* __restrict__ meta,
                                                        const uint32_t* __restrict__ totals,
                                                        const uint32_t* __restrict__ group_start,
                                                        const GridBwdBin* __restrict__ bins,
                                                        const uint32_t* __restrict__ run_pre,
                                                        const uint32_t* __restrict__ run_addr,
                                                        const uint32_t* __restrict__ recs,
                                                        long long* __restrict__ partial, float* __restrict__ grads,
                                                        int overwrite) {
    constexpr int RW = RecWords<F>::value;
    constexpr uint32_t kWaves = kBlock / 64;
    __shared__ unsigned long long acc[kTileFloats];
    __shared__ uint32_t s_pre[kMaxRuns + 1], s_addr[kMaxRuns];
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
    const GridBwdBin bin = bins[b];
    const uint32_t n_runs = bin.reps * n_qblocks;  // <= kMaxRuns (pass size)
    const uint32_t total = totals[b];
    for (uint32_t j = threadIdx.x; j < n_runs; j += kBlock) {
        s_pre[j] = run_pre[(size_t)bin.count_row * n_qblocks + j];
        s_addr[j] = run_addr[(size_t)bin.count_row * n_qblocks + j];
    }
    if (threadIdx.x == 0) s_pre[n_runs] = total;
    const uint32_t begin = (gi - group_start[b]) * kGroupRecs;
    const uint32_t end = min(total, begin + kGroupRecs);
    const uint32_t mbits = meta[3];
    const bool finite = finite_bits(mbits);
    const int S = finite ? fixed_shift(mbits) : 0;
    for (uint32_t i = threadIdx.x; i < kTileFloats; i += kBlock) acc[i] = 0ull;
    __syncthreads();
    // this wave's part of the group: [w0, w1), a multiple of 64 records long
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t per_wave = (end - begin + 64 * kWaves - 1) / (64 * kWaves) * 64;
    const uint32_t w0 = min(end, begin + wave * per_wave), w1 = min(end, w0 + per_wave);
    // the run of record p (p < total): l with s_pre[l] <= p < s_pre[l + 1];
    // binary search for the lane's first record, then forward steps
    uint32_t run = 0;
    {
        // the lane's first record: the first that load_batch asks for (later ones only grow)
        const uint32_t p = w0 < w1 ? min(w0 + lane, w1 - 1) : end - 1;
        uint32_t l = 0, h = n_runs;
        while (h - l > 1) {
            const uint32_t m = (l + h) >> 1;
            if (s_pre[m] <= p) l = m;
            else h = m;
        }
        run = l;
    }
    auto slot_of = [&](uint32_t p) -> size_t {
        while (s_pre[run + 1] <= p) ++run;  // ends: s_pre[n_runs] = total > p
        return (size_t)s_addr[run] + (p - s_pre[run]);
    };
    // software-pipelined: the next batch's loads are issued before this
    // batch's LDS adds (lanes past the end load the last record and add 0)
    uint32_t cur[kInFlight][RW] = {}, nxt[kInFlight][RW] = {};  // nxt stays unread after the last batch
    auto load_batch = [&](uint32_t i0, uint32_t (&dst)[kInFlight][RW]) {
        size_t at[kInFlight];
#pragma unroll
        for (uint32_t k = 0; k < kInFlight; ++k) at[k] = slot_of(min(i0 + k * 64 + lane, w1 - 1));
#pragma unroll
        for (uint32_t k = 0; k < kInFlight; ++k) load_rec<RW>(recs + at[k] * RW, dst[k]);
    };
    auto fixed = [&](float c) { return to_fixed(c, S); };
    if (w0 < w1) load_batch(w0, cur);
    for (uint32_t i0 = w0; i0 < w1; i0 += 64 * kInFlight) {
        const uint32_t i1 = i0 + 64 * kInFlight;
        if (i1 < w1) load_batch(i1, nxt);
#pragma unroll
        for (uint32_t k = 0; k < kInFlight; ++k) {
            const bool valid = i0 + k * 64 + lane < w1;
            const uint32_t key = cur[k][0];
            const bool edge = (key & kEdgeFlag) != 0;
            const uint32_t e0 = key & 0x1fffu, e1 = (key >> 13) & 0x1fffu;
            const float wt = __uint_as_float(cur[k][1]);
            const float w0f = edge ? 1.0f - wt : wt;  // an edge's weight word is its x fraction
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const float a = __uint_as_float(cur[k][2 + f]);
                atomicAdd(&acc[e0 * F + f], valid ? fixed(w0f * a) : 0ull);
                if (edge) atomicAdd(&acc[e1 * F + f], valid ? fixed(wt * a) : 0ull);
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kInFlight; ++k)
#pragma unroll
            for (int j = 0; j < RW; ++j) cur[k][j] = nxt[k][j];
    }
    __syncthreads();
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

// Pass 4: tiles with several groups, and (overwrite) tiles with none: the
// meta[2] tiles of apply_list, kTileFloats / kBlock work items each, taken
// by a persistent grid.
__global__ __launch_bounds__(kBlock) void apply_kernel(const GridBwdBin* __restrict__ bins,
                                                       const uint32_t* __restrict__ group_start,
                                                       const uint32_t* __restrict__ meta,
                                                       const uint32_t* __restrict__ apply_list,
                                                       const long long* __restrict__ partial,
                                                       float* __restrict__ grads, uint32_t F, int overwrite) {
    constexpr uint32_t kItems = kTileFloats / kBlock;
    const uint32_t n_items = meta[2] * kItems;
    const uint32_t mbits = meta[3];
    for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const uint32_t b = apply_list[item / kItems];
        const uint32_t c0 = group_start[b], c1 = group_start[b + 1];
        const uint32_t e = (item % kItems) * kBlock + threadIdx.x;
        const GridBwdBin bin = bins[b];
        if (e >= bin.n_entries * F) continue;
        float* dst = grads + bin.float_offset + e;
        if (c1 == c0) {
            *dst = 0.0f;
            continue;
        }
        long long sum = 0;
        for (uint32_t c = c0; c < c1; ++c) sum += partial[(size_t)c * kTileFloats + e];
        const float g = finite_bits(mbits) ? ldexpf((float)sum, -fixed_shift(mbits)) : __uint_as_float(0x7fc00000u);
        if (overwrite) *dst = g;
        else *dst += g;
    }
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
    for (const Table& t : tables) max_reps_ = std::max(max_reps_, t.reps);
    host_ = plan.release();
    supported_ = true;
}

GridBackward::~GridBackward() { delete host_; }

uint64_t GridBackward::max_groups(uint32_t n) const {
    // sum over tiles of ceil(t / G) <= R / G + tiles
    return (uint64_t)n * corners_per_query_ / kGroupRecs + n_bins_;
}

uint32_t GridBackward::queries_per_pass() const {
    // at most 2^27 records per pass (an entry then receives at most 2^27 /
    // (records per query and level) <= 2^kFixedHeadroom contributions), and at
    // most kMaxRuns runs (repetitions x query blocks) per tile
    const uint64_t cap = (1ull << 27) / std::max<uint32_t>(corners_per_query_, 1);
    const uint64_t runs_cap = (uint64_t)(kMaxRuns / std::max<uint32_t>(max_reps_, 1)) * kQueriesPerBlock;
    return (uint32_t)std::max<uint64_t>(kQueriesPerBlock,
                                        std::min(cap, runs_cap) / kQueriesPerBlock * kQueriesPerBlock);
}

void GridBackward::prepare(hipStream_t s, uint32_t N) {
    const uint32_t n = std::min(N, queries_per_pass());
    const uint32_t nqb = div_round_up(n, kQueriesPerBlock);
    counts_.ensure(sizeof(uint32_t) * (size_t)count_rows_ * nqb);
    run_addr_.ensure(sizeof(uint32_t) * (size_t)count_rows_ * nqb);
    totals_.ensure(sizeof(uint32_t) * n_bins_);
    chunk_start_.ensure(sizeof(uint32_t) * (n_bins_ + 1));
    apply_list_.ensure(sizeof(uint32_t) * n_bins_);
    if (!meta_.p) {
        meta_.ensure(sizeof(uint32_t) * 4);
        hipLaunchKernelGGL(reset_meta_kernel, dim3(1), dim3(64), 0, s, meta_.as<uint32_t>());
        NLOD_LAUNCH_CHECK();
    }
    const size_t rec_words = F_ + 2 <= 4 ? 4 : F_ + 2 <= 8 ? 8 : 16;
    // every (row, query block) has a region for its largest record count
    records_.ensure(sizeof(uint32_t) * (size_t)n_rows_ * nqb * kQueriesPerBlock * 8 * rec_words);
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
    prepare(s, N);
    for (uint32_t q0 = 0; q0 < N; q0 += pass) {
        const uint32_t n = std::min(pass, N - q0);
        const uint32_t nqb = div_round_up(n, kQueriesPerBlock);
        const uint64_t n_groups = max_groups(n);
        const float* in_c = in + (size_t)q0 * n_input_dims_;
        // d_enc is SoA [K][N]: the pass's columns start at q0 within each row of stride N
        const float* denc_c = d_enc + q0;
        by_F(F_, [&](auto Fc) {
            constexpr int F = decltype(Fc)::value;
            hipLaunchKernelGGL(bin_kernel<F>, dim3(nqb, n_rows_), dim3(kBlock), 0, s, n, N, nqb, n_input_dims_,
                               d_rows_.as<GridBwdRow>(), in_c, denc_c, counts_.as<uint32_t>(),
                               run_addr_.as<uint32_t>(), records_.as<uint32_t>(), meta_.as<uint32_t>());
        });
        NLOD_LAUNCH_CHECK();
        hipLaunchKernelGGL(scan_rows_kernel, dim3(n_bins_), dim3(kBlock), 0, s, nqb, d_bins_.as<GridBwdBin>(),
                           counts_.as<uint32_t>(), totals_.as<uint32_t>());
        NLOD_LAUNCH_CHECK();
        const int ow = (overwrite && q0 == 0) ? 1 : 0;
        hipLaunchKernelGGL(scan_bins_kernel, dim3(1), dim3(kBlock), 0, s, n_bins_, kGroupRecs, ow,
                           totals_.as<uint32_t>(), chunk_start_.as<uint32_t>(), apply_list_.as<uint32_t>(),
                           meta_.as<uint32_t>());
        NLOD_LAUNCH_CHECK();
        by_F(F_, [&](auto Fc) {
            constexpr int F = decltype(Fc)::value;
            hipLaunchKernelGGL(reduce_kernel<F>, dim3((uint32_t)n_groups), dim3(kBlock), 0, s, n_bins_, nqb,
                               meta_.as<uint32_t>(), totals_.as<uint32_t>(), chunk_start_.as<uint32_t>(),
                               d_bins_.as<GridBwdBin>(), counts_.as<uint32_t>(), run_addr_.as<uint32_t>(),
                               records_.as<uint32_t>(), partial_.as<long long>(), grads, ow);
        });
        NLOD_LAUNCH_CHECK();
        const uint32_t apply_blocks =
            std::min<uint32_t>(n_bins_ * (kTileFloats / kBlock), (uint32_t)cu_count() * 8);
        hipLaunchKernelGGL(apply_kernel, dim3(apply_blocks), dim3(kBlock), 0, s, d_bins_.as<GridBwdBin>(),
                           chunk_start_.as<uint32_t>(), meta_.as<uint32_t>(), apply_list_.as<uint32_t>(),
                           partial_.as<long long>(), grads, F_, ow);
        NLOD_LAUNCH_CHECK();
    }
}

}  // namespace nlod
