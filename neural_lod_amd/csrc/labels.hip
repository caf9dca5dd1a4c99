// Visibility training samples from a triangle mesh: rays that start in an
// occupied voxel of a res^3 LoD grid, labelled 1 when the mesh is hit before
// the ray leaves that voxel.
//
// The label is the visibility ray query of the reference renderer
// (rptr/vulkan/raygen.rgen:690-730: trace from the origin, take the nearest
// hit, occlusion = 1 when the hit point lies inside the voxel box).  The
// sampler is the one of the visibility training configuration: voxel uniform
// over the occupied voxels, origin uniform inside it, direction uniform on the
// sphere (a simplification of the reference's dilated-box sampler,
// rptr/cuda/neural-lod/neural_lod_learning_kernels.cuh:364-569).  The grid is
// the cubic mesh box at res^3 (neural_lod_learning_init.cu:137-172), and the
// occupied voxels are those of the surface voxelization.
//
// Exactly specified so that the CPU oracle reproduces every bit:
//   rng      per sample i: state = seed ^ (i * 0x9E3779B97F4A7C15), splitmix64
//            steps; u = (x >> 40) * 2^-24
//   voxel    slot = ((x >> 32) * n_occupied) >> 32 of the next draw; occupied
//            voxels in ascending linear index x + y res + z res^2
//   origin   o_k = bmin_k + ((float)v_k + u_k) * unit_k
//   dir      rejection in [-1, 1]^3 until 0 < s = a^2 + b^2 + c^2 <= 1 (at
//            most 64 draws, else (0, 0, 1)), then (a, b, c) / sqrtf(s)
//   hit      Moller-Trumbore against the voxel's triangles (those whose clamped
//            voxel range contains it, as the voxelizer computes it); nearest
//            t >= 0; label 1 when |p_k - c_k| <= unit_k / 2 for every axis
//            (p = o + t d, c the voxel centre)
//   inputs   [(o - bmin) / (bmax - bmin), (d + 1) / 2]
// Every float operation is a separately rounded IEEE single operation (no
// contraction); division and square root are correctly rounded.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "kernels.h"
#include "nlod.h"
#include "nlod_common.h"

namespace nlod {
namespace {

constexpr int kLabelThreads = 256;
constexpr int kMaxDirDraws = 64;  // then (0, 0, 1): probability 0.48^64

struct LabelGrid {
    float bmin[3], bmax[3], unit[3];
    uint32_t res, n_occupied;
};

__device__ __forceinline__ uint64_t splitmix64(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float uniform01(uint64_t& s) { return (float)(splitmix64(s) >> 40) * 0x1p-24f; }

struct v3 { float x, y, z; };
__device__ __forceinline__ v3 sub(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot(v3 a, v3 b) {
    const float t0 = a.x * b.x, t1 = a.y * b.y, t2 = a.z * b.z;
    return (t0 + t1) + t2;
}
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}

// Moller-Trumbore: t of the hit, or a negative value / NaN for none.
__device__ __forceinline__ float ray_triangle(v3 o, v3 d, const float* tri) {
    const v3 p0 = {tri[0], tri[1], tri[2]}, p1 = {tri[3], tri[4], tri[5]}, p2 = {tri[6], tri[7], tri[8]};
    const v3 e1 = sub(p1, p0), e2 = sub(p2, p0);
    const v3 pv = cross(d, e2);
    const float det = dot(e1, pv);
    if (det == 0.0f) return -1.0f;
    const float inv = 1.0f / det;
    const v3 s = sub(o, p0);
    const float u = dot(s, pv) * inv;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    const v3 q = cross(s, e1);
    const float w = dot(d, q) * inv;
    if (w < 0.0f || u + w > 1.0f) return -1.0f;
    return dot(e2, q) * inv;
}

__global__ __launch_bounds__(kLabelThreads) void visibility_samples_kernel(
    LabelGrid g, const float* __restrict__ tris, const uint32_t* __restrict__ occupied,
    const uint32_t* __restrict__ list_start, const uint32_t* __restrict__ list, uint64_t seed, uint64_t first,
    uint32_t n, float* __restrict__ inputs, float* __restrict__ labels) {
    const uint32_t i = blockIdx.x * kLabelThreads + threadIdx.x;
    if (i >= n) return;
    uint64_t st = seed ^ ((first + i) * 0x9E3779B97F4A7C15ull);
    const uint32_t slot = (uint32_t)(((splitmix64(st) >> 32) * (uint64_t)g.n_occupied) >> 32);
    const uint32_t lin = occupied[slot];
    const uint32_t vx = lin % g.res, vy = (lin / g.res) % g.res, vz = lin / (g.res * g.res);
    const float ux = uniform01(st), uy = uniform01(st), uz = uniform01(st);
    const v3 o = {g.bmin[0] + ((float)vx + ux) * g.unit[0], g.bmin[1] + ((float)vy + uy) * g.unit[1],
                  g.bmin[2] + ((float)vz + uz) * g.unit[2]};
    float a = 0.0f, b = 0.0f, c = 1.0f, s = 1.0f;
    for (int k = 0; k < kMaxDirDraws; ++k) {  // 1.9 draws of three expected
        const float a1 = uniform01(st) * 2.0f - 1.0f;
        const float b1 = uniform01(st) * 2.0f - 1.0f;
        const float c1 = uniform01(st) * 2.0f - 1.0f;
        const float s1 = (a1 * a1 + b1 * b1) + c1 * c1;
        if (s1 > 0.0f && s1 <= 1.0f) {
            a = a1; b = b1; c = c1; s = s1;
            break;
        }
    }
    const float r = sqrtf(s);
    const v3 d = {a / r, b / r, c / r};

    float best = INFINITY;
    for (uint32_t k = list_start[slot]; k < list_start[slot + 1]; ++k) {
        const float t = ray_triangle(o, d, tris + (size_t)list[k] * 9);
        if (t >= 0.0f && t < best) best = t;
    }
    float label = 0.0f;
    if (best < INFINITY) {
        const v3 p = {o.x + best * d.x, o.y + best * d.y, o.z + best * d.z};
        const v3 cc = {g.bmin[0] + ((float)vx + 0.5f) * g.unit[0], g.bmin[1] + ((float)vy + 0.5f) * g.unit[1],
                       g.bmin[2] + ((float)vz + 0.5f) * g.unit[2]};
        const bool inside = fabsf(p.x - cc.x) <= 0.5f * g.unit[0] && fabsf(p.y - cc.y) <= 0.5f * g.unit[1] &&
                            fabsf(p.z - cc.z) <= 0.5f * g.unit[2];
        label = inside ? 1.0f : 0.0f;
    }
    float* in = inputs + (size_t)i * 6;
    in[0] = (o.x - g.bmin[0]) / (g.bmax[0] - g.bmin[0]);
    in[1] = (o.y - g.bmin[1]) / (g.bmax[1] - g.bmin[1]);
    in[2] = (o.z - g.bmin[2]) / (g.bmax[2] - g.bmin[2]);
    in[3] = (d.x + 1.0f) * 0.5f;
    in[4] = (d.y + 1.0f) * 0.5f;
    in[5] = (d.z + 1.0f) * 0.5f;
    labels[i] = label;
}

// Clamped voxel range of a triangle on one axis, as the surface voxelizer
// computes it (cpu_voxelizer.cpp:70-84): (int)clamp((min v - bmin) / unit, 0, G-1).
// glm min(x, y) = (y < x) ? y : x, max(x, y) = (x < y) ? y : x
inline float gmin_h(float x, float y) { return (y < x) ? y : x; }
inline float gmax_h(float x, float y) { return (x < y) ? y : x; }

inline void axis_range(const float* tri, int k, float bmin, float unit, uint32_t G, int& lo, int& hi) {
    const float a = tri[k] - bmin, b = tri[3 + k] - bmin, c = tri[6 + k] - bmin;
    const float gm = (float)(G - 1);
    lo = (int)gmin_h(gmax_h(gmin_h(a, gmin_h(b, c)) / unit, 0.0f), gm);
    hi = (int)gmin_h(gmax_h(gmax_h(a, gmax_h(b, c)) / unit, 0.0f), gm);
}

}  // namespace

struct Labeler {
    LabelGrid grid{};
    const float* d_tris = nullptr;  // caller-owned, must outlive the labeler
    DevBuf occupied, list_start, list;
    uint64_t n_list = 0;
};

Labeler* labeler_create(hipStream_t s, const float* d_tris, uint64_t n_tri, uint32_t res) {
    NLOD_CHECK(d_tris != nullptr && n_tri > 0, "the mesh needs at least one triangle");
    NLOD_CHECK(res >= 1 && res <= 1024, "label grid resolution must be in [1, 1024]");
    NLOD_CHECK(n_tri < (1ull << 32), "at most 2^32 - 1 triangles");
    std::unique_ptr<Labeler> lb(new Labeler());
    float bmin[3], bmax[3];
    mesh_bbox(s, d_tris, n_tri, bmin, bmax);
    nlod_voxinfo info;
    voxinfo_from_bbox(bmin, bmax, res, 0, n_tri, &info);
    const uint64_t words = ((uint64_t)res * res * res + 31) / 32;
    DevBuf table;
    table.ensure(words * sizeof(uint32_t));
    voxelize(s, &info, d_tris, table.as<uint32_t>(), NLOD_VOX_SURFACE);
    std::vector<uint32_t> bits(words);
    std::vector<float> tris(n_tri * 9);
    NLOD_HIP(hipMemcpyAsync(bits.data(), table.p, words * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NLOD_HIP(hipMemcpyAsync(tris.data(), d_tris, tris.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    NLOD_HIP(hipStreamSynchronize(s));

    // occupied voxels in ascending linear index, and the slot of each voxel
    const uint64_t n_vox = (uint64_t)res * res * res;
    std::vector<uint32_t> occ;
    std::vector<int32_t> slot_of(n_vox, -1);
    for (uint64_t v = 0; v < n_vox; ++v)
        if (bits[v >> 5] & (1u << (31 - (v & 31)))) {
            slot_of[v] = (int32_t)occ.size();
            occ.push_back((uint32_t)v);
        }
    NLOD_CHECK(!occ.empty(), "the mesh occupies no voxel");
    // per occupied voxel: the triangles whose clamped voxel range contains it
    std::vector<uint32_t> count(occ.size() + 1, 0);
    auto for_voxels = [&](uint64_t t, auto&& f) {
        const float* tri = tris.data() + t * 9;
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(tri[k])) return;
        int lo[3], hi[3];
        for (int k = 0; k < 3; ++k) axis_range(tri, k, info.bbox_min[k], info.unit[k], res, lo[k], hi[k]);
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    const int32_t sl = slot_of[(uint64_t)x + (uint64_t)y * res + (uint64_t)z * res * res];
                    if (sl >= 0) f((uint32_t)sl);
                }
    };
    for (uint64_t t = 0; t < n_tri; ++t) for_voxels(t, [&](uint32_t sl) { ++count[sl + 1]; });
    for (size_t k = 1; k < count.size(); ++k) count[k] += count[k - 1];
    std::vector<uint32_t> fill(count.begin(), count.end() - 1), lst(count.back());
    for (uint64_t t = 0; t < n_tri; ++t) for_voxels(t, [&](uint32_t sl) { lst[fill[sl]++] = (uint32_t)t; });

    for (int k = 0; k < 3; ++k) {
        lb->grid.bmin[k] = info.bbox_min[k];
        lb->grid.bmax[k] = info.bbox_max[k];
        lb->grid.unit[k] = info.unit[k];
    }
    lb->grid.res = res;
    lb->grid.n_occupied = (uint32_t)occ.size();
    lb->d_tris = d_tris;
    lb->n_list = lst.size();
    lb->occupied.ensure(occ.size() * sizeof(uint32_t));
    lb->list_start.ensure(count.size() * sizeof(uint32_t));
    lb->list.ensure(std::max<size_t>(lst.size(), 1) * sizeof(uint32_t));
    NLOD_HIP(hipMemcpyAsync(lb->occupied.p, occ.data(), occ.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    NLOD_HIP(hipMemcpyAsync(lb->list_start.p, count.data(), count.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    if (!lst.empty())
        NLOD_HIP(hipMemcpyAsync(lb->list.p, lst.data(), lst.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    NLOD_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
    return lb.release();
}

void labeler_visibility(const Labeler& lb, hipStream_t s, uint64_t seed, uint64_t first, uint32_t n, float* inputs,
                        float* labels) {
    if (n == 0) return;
    NLOD_CHECK(inputs != nullptr && labels != nullptr, "output pointers are null");
    const uint32_t blocks = div_round_up(n, kLabelThreads);
    hipLaunchKernelGGL(visibility_samples_kernel, dim3(blocks), dim3(kLabelThreads), 0, s, lb.grid, lb.d_tris,
                       lb.occupied.as<uint32_t>(), lb.list_start.as<uint32_t>(), lb.list.as<uint32_t>(), seed, first,
                       n, inputs, labels);
    NLOD_LAUNCH_CHECK();
}

}  // namespace nlod

struct nlod_labeler {
    std::unique_ptr<nlod::Labeler> impl;
};

extern "C" {

int nlod_labeler_create(void* stream, const float* d_triangles, uint64_t n_triangles, uint32_t res,
                        nlod_labeler** out) {
    return nlod::guarded([&] {
        NLOD_CHECK(out != nullptr, "output pointer is null");
        *out = nullptr;
        std::unique_ptr<nlod_labeler> h(new nlod_labeler());
        h->impl.reset(nlod::labeler_create(nlod::as_stream(stream), d_triangles, n_triangles, res));
        *out = h.release();
    });
}

int nlod_labeler_destroy(nlod_labeler* labeler) {
    return nlod::guarded([&] { delete labeler; });
}

int nlod_labeler_info(const nlod_labeler* labeler, nlod_voxinfo* grid_out, uint32_t* n_occupied_out,
                      uint64_t* n_list_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(labeler != nullptr, "labeler is null");
        const nlod::LabelGrid& g = labeler->impl->grid;
        if (grid_out) {
            std::memset(grid_out, 0, sizeof(*grid_out));
            for (int k = 0; k < 3; ++k) {
                grid_out->bbox_min[k] = g.bmin[k];
                grid_out->bbox_max[k] = g.bmax[k];
                grid_out->gridsize[k] = g.res;
                grid_out->unit[k] = g.unit[k];
            }
        }
        if (n_occupied_out) *n_occupied_out = g.n_occupied;
        if (n_list_out) *n_list_out = labeler->impl->n_list;
    });
}

int nlod_labeler_visibility(const nlod_labeler* labeler, void* stream, uint64_t seed, uint64_t first, uint32_t n,
                            float* d_inputs, float* d_labels) {
    return nlod::guarded([&] {
        NLOD_CHECK(labeler != nullptr, "labeler is null");
        nlod::labeler_visibility(*labeler->impl, nlod::as_stream(stream), seed, first, n, d_inputs, d_labels);
    });
}

}  // extern "C"
