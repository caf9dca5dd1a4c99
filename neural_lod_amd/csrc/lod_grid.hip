// LoD occupancy pyramid and first-active-voxel ray marching.
//
// The reference builds a nanovdb grid per LoD from a point cloud downsampled
// by >> lod (rptr/cuda/neural-lod/neural_lod_learning_init.cu:201-295) and
// marches rays with nanovdb's TreeMarcher + DDA to the first active voxel
// that is not the previous one (first_active_voxel_intersection,
// neural_lod_learning_common.cuh:784-832, called by raymarch_impl,
// rptr/cuda/wavefront_neural_pt.cuh:89-141).  Here the pyramid is a stack of
// bit tables in the voxelizer's layout (bit 31 - i % 32 of word i / 32 for
// voxel i = x + y R + z R^2): level 0 is a voxelization, level l + 1 the OR of
// 2 x 2 x 2 blocks of level l.  The march is a 3D DDA (Amanatides & Woo 1987)
// over one level's bits, in that level's index space.
//
// March, specified so that the CPU oracle reproduces every bit (no FP
// contraction; division correctly rounded):
//   p0 = (o - bmin) / unit, q = d / unit per axis (unit of the level), so t is
//   the parameter of o + t d.  Clip to [0, R]^3 by slabs (an axis with q = 0
//   must have 0 <= p0 <= R), t from max(t_min, slab entry).  Start voxel
//   v = clamp(floor(p0 + t q), 0, R - 1).  Each step: if v is occupied and not
//   `prev`, its box [v, v + 1] by slabs gives t_near <= t_far -> hit; else
//   advance the axis whose next boundary ((v + (q > 0)) - p0) / q is nearest
//   (x before y before z on ties), stop past the grid exit or the grid edge;
//   at most 3 R + 3 steps.
//
// Per-voxel visibility thresholds (update_grid_with_thresholds,
// rptr/cuda/neural-lod/neural_lod_learning.cu:50-121): the reference stores
// one float per active voxel of a level, in increasing Morton order
// (x in bit 0, y in bit 1, z in bit 2; neural_lod_learning.h:519-532,
// neural_lod_learning_init.cu:288-295), clamped to [min, max], and the march
// reports the hit voxel's value (0.5 where none were set,
// neural_lod_learning_init.cu:242).  Here a level keeps a dense float table of
// R^3 values; the Morton rank of each active voxel comes from per-word
// popcounts over the Morton-ordered bits and a two-level exclusive scan.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "nlod.h"
#include "nlod_common.h"

namespace nlod {
namespace {

constexpr int kMarchThreads = 256;
constexpr int kMaxLodLevels = 10;
constexpr uint32_t kNoVoxel = 0xffffffffu;

// One output voxel per lane; a wave covers 64 consecutive voxels = 2 words.
__global__ __launch_bounds__(256) void downsample_kernel(const uint32_t* __restrict__ child, uint32_t Rc,
                                                         uint32_t* __restrict__ parent, uint32_t Rp) {
    const uint64_t n_vox = (uint64_t)Rp * Rp * Rp;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool occ = false;
    if (i < n_vox) {
        const uint32_t x = (uint32_t)(i % Rp), y = (uint32_t)((i / Rp) % Rp), z = (uint32_t)(i / ((uint64_t)Rp * Rp));
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const uint64_t ci = (uint64_t)(2 * x + (c & 1)) + (uint64_t)(2 * y + ((c >> 1) & 1)) * Rc +
                                (uint64_t)(2 * z + (c >> 2)) * Rc * Rc;
            occ = occ || ((child[ci >> 5] >> (31u - (uint32_t)(ci & 31u))) & 1u);
        }
    }
    const uint64_t mask = __ballot(occ);
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t w0 = (i - lane) >> 5;  // first of this wave's two words
    const uint64_t n_words = (n_vox + 31) / 32;
    if (lane == 0 && w0 < n_words) parent[w0] = __builtin_bitreverse32((uint32_t)mask);
    if (lane == 0 && w0 + 1 < n_words) parent[w0 + 1] = __builtin_bitreverse32((uint32_t)(mask >> 32));
}

struct MarchLevel {
    const uint32_t* bits;
    uint32_t res;
    float bmin[3], unit[3];
};

__device__ __forceinline__ bool occupied(const uint32_t* bits, uint32_t lin) {
    return (bits[lin >> 5] >> (31u - (lin & 31u))) & 1u;
}

constexpr float kDefaultThreshold = 0.5f;
constexpr int kRankBlock = 256;  // Morton words per block of the rank kernels

// Every third bit of a Morton code gathered into one axis coordinate
// (bits 0, 3, 6, ... -> x for the code itself, y for code >> 1, z for code >> 2).
__device__ __forceinline__ uint32_t compact3(uint32_t v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030c30c3u;
    v = (v | (v >> 4)) & 0x0300f00fu;
    v = (v | (v >> 8)) & 0xff0000ffu;
    v = (v | (v >> 16)) & 0x3ffu;
    return v;
}

// Morton word w covers Morton codes 32 w .. 32 w + 31 of a power-of-two level
// (R^3 codes, R^3 / 32 words).  Writes the word's occupancy mask (bit b for
// code 32 w + b), and per block the exclusive prefix of popcounts inside the
// block plus the block total.
__global__ __launch_bounds__(kRankBlock) void morton_rank_kernel(const uint32_t* __restrict__ bits, uint32_t R,
                                                                 uint32_t n_words, uint32_t* __restrict__ mask_out,
                                                                 uint32_t* __restrict__ local_out,
                                                                 uint32_t* __restrict__ block_total) {
    __shared__ uint32_t s_wave[kRankBlock / 64];
    const uint32_t w = blockIdx.x * kRankBlock + threadIdx.x;
    uint32_t mask = 0;
    if (w < n_words) {
        // R^3 < 32 (R = 1, 2): the only word holds fewer codes
        const uint32_t nb = (uint32_t)min<uint64_t>(32u, (uint64_t)R * R * R - 32ull * w);
        for (uint32_t b = 0; b < nb; ++b) {
            const uint32_t code = 32u * w + b;
            const uint32_t x = compact3(code), y = compact3(code >> 1), z = compact3(code >> 2);
            const uint64_t lin = (uint64_t)x + (uint64_t)y * R + (uint64_t)z * R * R;
            mask |= ((bits[lin >> 5] >> (31u - (uint32_t)(lin & 31u))) & 1u) << b;
        }
        mask_out[w] = mask;
    }
    // block exclusive scan of popcounts: wave inclusive scan, then wave offsets
    const uint32_t c = (uint32_t)__popc(mask);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t inc = c;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t k = 0; k < wave; ++k) off += s_wave[k];
    if (w < n_words) local_out[w] = off + inc - c;
    if (threadIdx.x == kRankBlock - 1) {
        uint32_t t = 0;
        for (uint32_t k = 0; k < kRankBlock / 64; ++k) t += s_wave[k];
        block_total[blockIdx.x] = t;
    }
}

// One block: block totals -> exclusive block offsets (in place) and the grand
// total in total_out[0].  Each lane scans a contiguous run of blocks.
__global__ __launch_bounds__(1024) void block_offsets_kernel(uint32_t* __restrict__ totals, uint32_t n_blocks,
                                                             uint64_t* __restrict__ total_out) {
    __shared__ uint64_t s_run[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n_blocks + 1023) / 1024;
    const uint32_t b0 = min(n_blocks, t * per), b1 = min(n_blocks, b0 + per);
    uint64_t sum = 0;
    for (uint32_t b = b0; b < b1; ++b) sum += totals[b];
    s_run[t] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint64_t o = t >= d ? s_run[t - d] : 0;
        __syncthreads();
        s_run[t] += o;
        __syncthreads();
    }
    uint64_t run = s_run[t] - sum;
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t v = totals[b];
        totals[b] = (uint32_t)run;
        run += v;
    }
    if (t == 1023) total_out[0] = s_run[1023];
}

__global__ void fill_kernel(float* __restrict__ out, uint64_t n, float v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v;
}

// Active voxel of Morton rank r takes clamp(thr[r], lo, hi) (the reference's
// glm::clamp: min(max(v, lo), hi)).
__global__ __launch_bounds__(kRankBlock) void scatter_thresholds_kernel(
    uint32_t R, uint32_t n_words, const uint32_t* __restrict__ masks, const uint32_t* __restrict__ local,
    const uint32_t* __restrict__ block_off, const float* __restrict__ thr, float lo, float hi,
    float* __restrict__ dense) {
    const uint32_t w = blockIdx.x * kRankBlock + threadIdx.x;
    if (w >= n_words) return;
    uint32_t mask = masks[w];
    uint32_t rank = block_off[blockIdx.x] + local[w];
    while (mask) {
        const uint32_t b = (uint32_t)__builtin_ctz(mask);
        mask &= mask - 1u;
        const uint32_t code = 32u * w + b;
        const uint32_t x = compact3(code), y = compact3(code >> 1), z = compact3(code >> 2);
        const uint64_t lin = (uint64_t)x + (uint64_t)y * R + (uint64_t)z * R * R;
        dense[lin] = fminf(fmaxf(thr[rank], lo), hi);
        ++rank;
    }
}

__global__ __launch_bounds__(kMarchThreads) void march_kernel(MarchLevel g, uint32_t n, const float* __restrict__ orig,
                                                              const float* __restrict__ dirs,
                                                              const float* __restrict__ t_min_in,
                                                              const uint32_t* __restrict__ prev_in,
                                                              uint32_t* __restrict__ voxel_out,
                                                              float* __restrict__ t_near_out,
                                                              float* __restrict__ t_far_out,
                                                              const float* __restrict__ thr_table,
                                                              float* __restrict__ thr_out) {
    const uint32_t i = blockIdx.x * kMarchThreads + threadIdx.x;
    if (i >= n) return;
    const float R = (float)g.res;
    float p0[3], q[3];
    for (int a = 0; a < 3; ++a) {
        p0[a] = (orig[(size_t)i * 3 + a] - g.bmin[a]) / g.unit[a];
        q[a] = dirs[(size_t)i * 3 + a] / g.unit[a];
    }
    const float t_min = t_min_in ? t_min_in[i] : 0.0f;
    const uint32_t prev = prev_in ? prev_in[i] : kNoVoxel;
    uint32_t found = kNoVoxel;
    float hit_near = -1.0f, hit_far = -1.0f;
    float tn = t_min, tf = INFINITY;
    bool miss = false;
    for (int a = 0; a < 3; ++a) {
        if (q[a] == 0.0f) {
            if (!(p0[a] >= 0.0f && p0[a] <= R)) miss = true;
            continue;
        }
        float t0 = (0.0f - p0[a]) / q[a], t1 = (R - p0[a]) / q[a];
        if (t1 < t0) { const float tmp = t0; t0 = t1; t1 = tmp; }
        if (t0 > tn) tn = t0;
        if (t1 < tf) tf = t1;
    }
    if (!miss && tn <= tf) {
        int v[3];
        for (int a = 0; a < 3; ++a) {
            const float c = floorf(p0[a] + tn * q[a]);
            v[a] = c < 0.0f ? 0 : (c > R - 1.0f ? (int)g.res - 1 : (int)c);
        }
        const int max_steps = 3 * (int)g.res + 3;
        for (int step = 0; step < max_steps; ++step) {
            const uint32_t lin = (uint32_t)v[0] + (uint32_t)v[1] * g.res + (uint32_t)v[2] * g.res * g.res;
            if (lin != prev && occupied(g.bits, lin)) {
                float bn = -INFINITY, bf = INFINITY;
                for (int a = 0; a < 3; ++a) {
                    if (q[a] == 0.0f) continue;
                    float t0 = ((float)v[a] - p0[a]) / q[a], t1 = ((float)(v[a] + 1) - p0[a]) / q[a];
                    if (t1 < t0) { const float tmp = t0; t0 = t1; t1 = tmp; }
                    if (t0 > bn) bn = t0;
                    if (t1 < bf) bf = t1;
                }
                if (bn <= bf) {
                    found = lin;
                    hit_near = bn;
                    hit_far = bf;
                    break;
                }
            }
            int axis = -1;
            float best = INFINITY;
            for (int a = 0; a < 3; ++a) {
                if (q[a] == 0.0f) continue;
                const float tb = ((float)(v[a] + (q[a] > 0.0f ? 1 : 0)) - p0[a]) / q[a];
                if (tb < best) {
                    best = tb;
                    axis = a;
                }
            }
            if (axis < 0 || best > tf) break;
            v[axis] += q[axis] > 0.0f ? 1 : -1;
            if (v[axis] < 0 || v[axis] >= (int)g.res) break;
        }
    }
    voxel_out[i] = found;
    t_near_out[i] = hit_near;
    t_far_out[i] = hit_far;
    if (thr_out) thr_out[i] = found == kNoVoxel ? -1.0f : (thr_table ? thr_table[found] : kDefaultThreshold);
}

}  // namespace

struct LodGrid {
    float bmin[3], bmax[3];
    uint32_t n_lods = 0;
    uint32_t res[kMaxLodLevels] = {};
    float unit[kMaxLodLevels][3] = {};
    DevBuf bits[kMaxLodLevels];
    DevBuf thr[kMaxLodLevels];  // dense per-voxel thresholds, once set
};

namespace {

// Morton ranks of a level's active voxels: masks, in-block prefixes and block
// offsets in the buffers; returns the number of active voxels (synchronizes).
uint64_t morton_ranks(const LodGrid& g, hipStream_t s, uint32_t lod, DevBuf& masks, DevBuf& local, DevBuf& blocks,
                      DevBuf& total) {
    const uint32_t R = g.res[lod];
    NLOD_CHECK((R & (R - 1)) == 0 && R <= 1024,
               "per-voxel thresholds need a power-of-two level resolution of at most 1024 (Morton order)");
    const uint64_t n_vox = (uint64_t)R * R * R;
    const uint32_t n_words = (uint32_t)((n_vox + 31) / 32);
    const uint32_t n_blocks = div_round_up(n_words, (uint32_t)kRankBlock);
    masks.ensure(sizeof(uint32_t) * n_words);
    local.ensure(sizeof(uint32_t) * n_words);
    blocks.ensure(sizeof(uint32_t) * n_blocks);
    total.ensure(sizeof(uint64_t));
    hipLaunchKernelGGL(morton_rank_kernel, dim3(n_blocks), dim3(kRankBlock), 0, s, g.bits[lod].as<uint32_t>(), R,
                       n_words, masks.as<uint32_t>(), local.as<uint32_t>(), blocks.as<uint32_t>());
    NLOD_LAUNCH_CHECK();
    hipLaunchKernelGGL(block_offsets_kernel, dim3(1), dim3(1024), 0, s, blocks.as<uint32_t>(), n_blocks,
                       total.as<uint64_t>());
    NLOD_LAUNCH_CHECK();
    uint64_t count = 0;
    NLOD_HIP(hipMemcpyAsync(&count, total.p, sizeof(count), hipMemcpyDeviceToHost, s));
    NLOD_HIP(hipStreamSynchronize(s));
    return count;
}

}  // namespace

uint64_t lod_grid_active_voxels(const LodGrid& g, hipStream_t s, uint32_t lod) {
    NLOD_CHECK(lod < g.n_lods, "lod out of range");
    DevBuf masks, local, blocks, total;
    return morton_ranks(g, s, lod, masks, local, blocks, total);
}

void lod_grid_set_thresholds(LodGrid& g, hipStream_t s, uint32_t lod, uint64_t n, const float* thr, float lo,
                             float hi) {
    NLOD_CHECK(lod < g.n_lods, "lod out of range");
    NLOD_CHECK(n == 0 || thr != nullptr, "null buffer");
    NLOD_CHECK(lo <= hi, "threshold range is empty (min > max)");
    DevBuf masks, local, blocks, total;
    const uint64_t active = morton_ranks(g, s, lod, masks, local, blocks, total);
    if (n != active)
        fail(NLOD_ERR_INVALID_ARGUMENT, "expected one threshold per active voxel of the level (" +
                                            std::to_string(active) + "), got " + std::to_string(n));
    const uint32_t R = g.res[lod];
    const uint64_t n_vox = (uint64_t)R * R * R;
    const uint32_t n_words = (uint32_t)((n_vox + 31) / 32);
    g.thr[lod].ensure(sizeof(float) * n_vox);
    hipLaunchKernelGGL(fill_kernel, dim3((uint32_t)((n_vox + 255) / 256)), dim3(256), 0, s, g.thr[lod].as<float>(),
                       n_vox, kDefaultThreshold);
    NLOD_LAUNCH_CHECK();
    hipLaunchKernelGGL(scatter_thresholds_kernel, dim3(div_round_up(n_words, (uint32_t)kRankBlock)), dim3(kRankBlock),
                       0, s, R, n_words, masks.as<uint32_t>(), local.as<uint32_t>(), blocks.as<uint32_t>(), thr, lo,
                       hi, g.thr[lod].as<float>());
    NLOD_LAUNCH_CHECK();
    NLOD_HIP(hipStreamSynchronize(s));  // the rank buffers are released on return
}

LodGrid* lod_grid_create(hipStream_t s, const nlod_voxinfo* info, const uint32_t* d_table, uint32_t n_lods) {
    NLOD_CHECK(info != nullptr && d_table != nullptr, "null argument");
    const uint32_t R = info->gridsize[0];
    NLOD_CHECK(R > 0 && info->gridsize[1] == R && info->gridsize[2] == R, "the grid must be cubic");
    NLOD_CHECK(R <= 4096, "grid resolution above 4096 is not supported");
    NLOD_CHECK(n_lods >= 1 && n_lods <= (uint32_t)kMaxLodLevels, "n_lods must be in [1, 10]");
    NLOD_CHECK(R % (1u << (n_lods - 1)) == 0, "the resolution must be divisible by 2^(n_lods - 1)");
    std::unique_ptr<LodGrid> g(new LodGrid());
    g->n_lods = n_lods;
    for (int k = 0; k < 3; ++k) {
        g->bmin[k] = info->bbox_min[k];
        g->bmax[k] = info->bbox_max[k];
    }
    for (uint32_t l = 0; l < n_lods; ++l) {
        const uint32_t Rl = R >> l;
        g->res[l] = Rl;
        // level 0 keeps the voxelizer's unit; coarser units from the box, as voxinfo computes units
        for (int k = 0; k < 3; ++k)
            g->unit[l][k] = l == 0 ? info->unit[k] : (info->bbox_max[k] - info->bbox_min[k]) / (float)Rl;
        const uint64_t words = ((uint64_t)Rl * Rl * Rl + 31) / 32;
        g->bits[l].ensure(words * sizeof(uint32_t));
        if (l == 0) {
            NLOD_HIP(hipMemcpyAsync(g->bits[0].p, d_table, words * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        } else {
            const uint64_t n_vox = (uint64_t)Rl * Rl * Rl;
            const uint32_t blocks = (uint32_t)((n_vox + 255) / 256);
            hipLaunchKernelGGL(downsample_kernel, dim3(blocks), dim3(256), 0, s, g->bits[l - 1].as<uint32_t>(),
                               g->res[l - 1], g->bits[l].as<uint32_t>(), Rl);
            NLOD_LAUNCH_CHECK();
        }
    }
    return g.release();
}

void lod_grid_march(const LodGrid& g, hipStream_t s, uint32_t lod, uint32_t n, const float* origins,
                    const float* dirs, const float* t_min, const uint32_t* prev, uint32_t* voxel_out, float* t_near,
                    float* t_far, float* thr_out) {
    NLOD_CHECK(lod < g.n_lods, "lod out of range");
    if (n == 0) return;
    NLOD_CHECK(origins && dirs && voxel_out && t_near && t_far, "null buffer");
    MarchLevel m{g.bits[lod].as<uint32_t>(), g.res[lod], {g.bmin[0], g.bmin[1], g.bmin[2]},
                 {g.unit[lod][0], g.unit[lod][1], g.unit[lod][2]}};
    hipLaunchKernelGGL(march_kernel, dim3(div_round_up(n, kMarchThreads)), dim3(kMarchThreads), 0, s, m, n, origins,
                       dirs, t_min, prev, voxel_out, t_near, t_far, g.thr[lod].as<float>(), thr_out);
    NLOD_LAUNCH_CHECK();
}

// What the path tracer reads of a level: bits, dense thresholds (null when
// unset), resolution, box, the level-0 voxel edge and resolution.
void lod_grid_tracer_view(const LodGrid& g, uint32_t lod, const uint32_t** bits, const float** thr, uint32_t* res,
                          float bmin[3], float bmax[3], float* voxel_size, uint32_t* grid_res) {
    NLOD_CHECK(lod < g.n_lods, "lod out of range");
    *bits = g.bits[lod].as<uint32_t>();
    *thr = g.thr[lod].as<float>();
    *res = g.res[lod];
    for (int k = 0; k < 3; ++k) {
        bmin[k] = g.bmin[k];
        bmax[k] = g.bmax[k];
    }
    *voxel_size = g.unit[0][0];
    *grid_res = g.res[0];
}

}  // namespace nlod

struct nlod_lod_grid {
    std::unique_ptr<nlod::LodGrid> impl;
};

namespace nlod {
const LodGrid& lod_grid_of(const nlod_lod_grid* g) {
    NLOD_CHECK(g && g->impl, "null grid");
    return *g->impl;
}
}  // namespace nlod

extern "C" {

int nlod_lod_grid_create(void* stream, const nlod_voxinfo* info, const uint32_t* d_table, uint32_t n_lods,
                         nlod_lod_grid** out) {
    return nlod::guarded([&] {
        NLOD_CHECK(out != nullptr, "output pointer is null");
        *out = nullptr;
        std::unique_ptr<nlod_lod_grid> h(new nlod_lod_grid());
        h->impl.reset(nlod::lod_grid_create(nlod::as_stream(stream), info, d_table, n_lods));
        *out = h.release();
    });
}

int nlod_lod_grid_destroy(nlod_lod_grid* grid) {
    return nlod::guarded([&] { delete grid; });
}

int nlod_lod_grid_level(const nlod_lod_grid* grid, uint32_t lod, const uint32_t** d_table_out,
                        nlod_voxinfo* info_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(grid != nullptr, "grid is null");
        const nlod::LodGrid& g = *grid->impl;
        NLOD_CHECK(lod < g.n_lods, "lod out of range");
        if (d_table_out) *d_table_out = g.bits[lod].as<uint32_t>();
        if (info_out) {
            std::memset(info_out, 0, sizeof(*info_out));
            for (int k = 0; k < 3; ++k) {
                info_out->bbox_min[k] = g.bmin[k];
                info_out->bbox_max[k] = g.bmax[k];
                info_out->gridsize[k] = g.res[lod];
                info_out->unit[k] = g.unit[lod][k];
            }
        }
    });
}

int nlod_lod_grid_march(const nlod_lod_grid* grid, void* stream, uint32_t lod, uint32_t n, const float* d_origins,
                        const float* d_dirs, const float* d_t_min, const uint32_t* d_prev_voxel,
                        uint32_t* d_voxel_out, float* d_t_near_out, float* d_t_far_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(grid != nullptr, "grid is null");
        nlod::lod_grid_march(*grid->impl, nlod::as_stream(stream), lod, n, d_origins, d_dirs, d_t_min, d_prev_voxel,
                             d_voxel_out, d_t_near_out, d_t_far_out, nullptr);
    });
}

int nlod_lod_grid_march_thresholds(const nlod_lod_grid* grid, void* stream, uint32_t lod, uint32_t n,
                                   const float* d_origins, const float* d_dirs, const float* d_t_min,
                                   const uint32_t* d_prev_voxel, uint32_t* d_voxel_out, float* d_t_near_out,
                                   float* d_t_far_out, float* d_threshold_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(grid != nullptr, "grid is null");
        NLOD_CHECK(n == 0 || d_threshold_out != nullptr, "null buffer");
        nlod::lod_grid_march(*grid->impl, nlod::as_stream(stream), lod, n, d_origins, d_dirs, d_t_min, d_prev_voxel,
                             d_voxel_out, d_t_near_out, d_t_far_out, d_threshold_out);
    });
}

int nlod_lod_grid_active_voxels(const nlod_lod_grid* grid, void* stream, uint32_t lod, uint64_t* count_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(grid != nullptr && count_out != nullptr, "null argument");
        *count_out = nlod::lod_grid_active_voxels(*grid->impl, nlod::as_stream(stream), lod);
    });
}

int nlod_lod_grid_set_thresholds(nlod_lod_grid* grid, void* stream, uint32_t lod, uint64_t n,
                                 const float* d_thresholds, float min_threshold, float max_threshold) {
    return nlod::guarded([&] {
        NLOD_CHECK(grid != nullptr, "grid is null");
        nlod::lod_grid_set_thresholds(*grid->impl, nlod::as_stream(stream), lod, n, d_thresholds, min_threshold,
                                      max_threshold);
    });
}

}  // extern "C"
