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
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
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

__global__ __launch_bounds__(kMarchThreads) void march_kernel(MarchLevel g, uint32_t n, const float* __restrict__ orig,
                                                              const float* __restrict__ dirs,
                                                              const float* __restrict__ t_min_in,
                                                              const uint32_t* __restrict__ prev_in,
                                                              uint32_t* __restrict__ voxel_out,
                                                              float* __restrict__ t_near_out,
                                                              float* __restrict__ t_far_out) {
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
}

}  // namespace

struct LodGrid {
    float bmin[3], bmax[3];
    uint32_t n_lods = 0;
    uint32_t res[kMaxLodLevels] = {};
    float unit[kMaxLodLevels][3] = {};
    DevBuf bits[kMaxLodLevels];
};

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
                    float* t_far) {
    NLOD_CHECK(lod < g.n_lods, "lod out of range");
    if (n == 0) return;
    NLOD_CHECK(origins && dirs && voxel_out && t_near && t_far, "null buffer");
    MarchLevel m{g.bits[lod].as<uint32_t>(), g.res[lod], {g.bmin[0], g.bmin[1], g.bmin[2]},
                 {g.unit[lod][0], g.unit[lod][1], g.unit[lod][2]}};
    hipLaunchKernelGGL(march_kernel, dim3(div_round_up(n, kMarchThreads)), dim3(kMarchThreads), 0, s, m, n, origins,
                       dirs, t_min, prev, voxel_out, t_near, t_far);
    NLOD_LAUNCH_CHECK();
}

}  // namespace nlod

struct nlod_lod_grid {
    std::unique_ptr<nlod::LodGrid> impl;
};

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
                             d_voxel_out, d_t_near_out, d_t_far_out);
    });
}

}  // extern "C"
