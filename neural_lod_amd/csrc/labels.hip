// Training samples from a triangle mesh for the two LoD networks.
//
// Visibility: rays that start in an occupied voxel, labelled 1 when the mesh
// is hit before the ray leaves that voxel -- the renderer's visibility query
// (rptr/vulkan/raygen.rgen:690-730: trace from the origin, nearest hit,
// occlusion 1 when the hit point lies inside the voxel box).
//
// Throughput: the voxel-local RGB throughput of the renderer's throughput
// query (raygen.rgen:505-688): a ray enters the voxel along -wo; at each hit
// inside the voxel a shadow ray along the fixed light direction wi adds
// T * albedo / pi * cos when it leaves the voxel unoccluded, then a cosine-
// weighted bounce multiplies T by the albedo; the path ends when it leaves the
// voxel, misses, or after max_depth hits.  A path whose first ray misses the
// voxel's geometry is invalid (valid_contribution = 0 there); the target is
// the mean of spp valid paths' throughputs (0 with valid = 0 when none is).
// Surfaces are two-sided Lambertian with a per-triangle albedo.
//
// The LoD pyramid: level l is the res >> l grid over the same cubic mesh box
// (neural_lod_learning_init.cu:137-172, 201-295: point cloud downsampled by
// >> lod); its occupancy is the OR of the level-0 surface voxelization over
// 2^l blocks.  Sampling (simplified from the reference's dilated-box sampler,
// rptr/cuda/neural-lod/neural_lod_learning_kernels.cuh:100-238, 364-569):
// level uniform, voxel uniform over the level's occupied voxels, directions
// uniform on the sphere.
//
// Exactly specified so that the CPU oracle reproduces every bit:
//   rng      per sample i: state = seed ^ (i * 0x9E3779B97F4A7C15), splitmix64
//            steps; u = (x >> 40) * 2^-24; an index in [0, m) is
//            ((x >> 32) * m) >> 32
//   sphere   rejection in [-1, 1]^3 until 0 < s <= 1 (64 draws at most, else
//            (0, 0, 1)), then (a, b, c) / sqrtf(s)
//   visib.   level 0; origin bmin + ((float)v + u) * unit per axis; label from
//            the nearest Moller-Trumbore hit t >= 0 among the voxel's
//            triangles: 1 when |p - centre| <= unit / 2 on every axis;
//            inputs [(o - bmin) / (bmax - bmin), (d + 1) / 2]
//   throughput  level = index draw over n_lods, voxel, wo, wi; per path: entry
//            by rejection of plane offsets (ox, oy) in [-U, U]^2 (U = largest
//            unit of the level) around centre + 2 U wo until the ray along -wo
//            meets the voxel box (64 draws at most, else the path is invalid),
//            then the path above with t_min = 1e-4 (1 + |origin|) after the
//            first ray, disk rejection for the cosine bounce in the Duff et
//            al. basis of the normal; inputs [(centre - bmin) / (bmax - bmin),
//            (wo + 1) / 2, (wi + 1) / 2, level]
//   albedo   component k of triangle t: 0.1 + 0.8 u, u from the splitmix64
//            stream seeded albedo_seed ^ (t * 0x9E3779B97F4A7C15), draw k
//   lists    per occupied voxel of a level: the finite triangles whose clamped
//            voxel range at that level's resolution contains it, as the
//            voxelizer computes ranges (cpu_voxelizer.cpp:70-84)
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
constexpr int kMaxDraws = 64;  // rejection loops: then a fixed fallback (probability < 0.75^64)
constexpr int kMaxLods = 8;
constexpr float kRayEps = 1e-4f;          // defaults.glsl:15 EPSILON
constexpr float kInvPi = 0.318309886183790671538f;
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

struct LevelDev {
    const uint32_t* occupied;
    const uint32_t* list_start;
    const uint32_t* list;
    uint32_t res, n_occupied;
    float unit[3];
};

struct LevelsDev {
    LevelDev l[kMaxLods];
    float bmin[3], bmax[3];
    uint32_t n_lods;
};

__device__ __forceinline__ uint64_t splitmix64(uint64_t& s) {
    s += kGolden;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float uniform01(uint64_t& s) { return (float)(splitmix64(s) >> 40) * 0x1p-24f; }
__device__ __forceinline__ uint32_t index_draw(uint64_t& s, uint32_t m) {
    return (uint32_t)(((splitmix64(s) >> 32) * (uint64_t)m) >> 32);
}

struct v3 { float x, y, z; };
__device__ __forceinline__ v3 sub(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot(v3 a, v3 b) {
    const float t0 = a.x * b.x, t1 = a.y * b.y, t2 = a.z * b.z;
    return (t0 + t1) + t2;
}
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ v3 along(v3 o, float t, v3 d) { return {o.x + t * d.x, o.y + t * d.y, o.z + t * d.z}; }

__device__ v3 sphere_dir(uint64_t& st) {
    float a = 0.0f, b = 0.0f, c = 1.0f, s = 1.0f;
    for (int k = 0; k < kMaxDraws; ++k) {
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
    return {a / r, b / r, c / r};
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

// Nearest hit with t >= t_min among the voxel's triangles (first in list order on ties).
__device__ __forceinline__ float nearest_hit(const LevelDev& L, uint32_t slot, const float* tris, v3 o, v3 d,
                                             float t_min, uint32_t& tri_out) {
    float best = INFINITY;
    for (uint32_t k = L.list_start[slot]; k < L.list_start[slot + 1]; ++k) {
        const uint32_t t = L.list[k];
        const float th = ray_triangle(o, d, tris + (size_t)t * 9);
        if (th >= t_min && th < best) {
            best = th;
            tri_out = t;
        }
    }
    return best;
}

struct Box {
    v3 lo, hi, c, h;
};

__device__ __forceinline__ Box voxel_box(const LevelsDev& g, const LevelDev& L, uint32_t vx, uint32_t vy,
                                         uint32_t vz) {
    Box b;
    b.lo = {g.bmin[0] + (float)vx * L.unit[0], g.bmin[1] + (float)vy * L.unit[1], g.bmin[2] + (float)vz * L.unit[2]};
    b.hi = {g.bmin[0] + (float)(vx + 1) * L.unit[0], g.bmin[1] + (float)(vy + 1) * L.unit[1],
            g.bmin[2] + (float)(vz + 1) * L.unit[2]};
    b.c = {g.bmin[0] + ((float)vx + 0.5f) * L.unit[0], g.bmin[1] + ((float)vy + 0.5f) * L.unit[1],
           g.bmin[2] + ((float)vz + 0.5f) * L.unit[2]};
    b.h = {0.5f * L.unit[0], 0.5f * L.unit[1], 0.5f * L.unit[2]};
    return b;
}

__device__ __forceinline__ bool in_box(const Box& b, v3 p) {
    return fabsf(p.x - b.c.x) <= b.h.x && fabsf(p.y - b.c.y) <= b.h.y && fabsf(p.z - b.c.z) <= b.h.z;
}

__device__ __forceinline__ void voxel_xyz(uint32_t lin, uint32_t res, uint32_t& x, uint32_t& y, uint32_t& z) {
    x = lin % res;
    y = (lin / res) % res;
    z = lin / (res * res);
}

__global__ __launch_bounds__(kLabelThreads) void visibility_samples_kernel(LevelsDev g, const float* __restrict__ tris,
                                                                           uint64_t seed, uint64_t first, uint32_t n,
                                                                           float* __restrict__ inputs,
                                                                           float* __restrict__ labels) {
    const uint32_t i = blockIdx.x * kLabelThreads + threadIdx.x;
    if (i >= n) return;
    const LevelDev& L = g.l[0];
    uint64_t st = seed ^ ((first + i) * kGolden);
    const uint32_t slot = index_draw(st, L.n_occupied);
    uint32_t vx, vy, vz;
    voxel_xyz(L.occupied[slot], L.res, vx, vy, vz);
    const float ux = uniform01(st), uy = uniform01(st), uz = uniform01(st);
    const v3 o = {g.bmin[0] + ((float)vx + ux) * L.unit[0], g.bmin[1] + ((float)vy + uy) * L.unit[1],
                  g.bmin[2] + ((float)vz + uz) * L.unit[2]};
    const v3 d = sphere_dir(st);
    uint32_t tri = 0;
    const float best = nearest_hit(L, slot, tris, o, d, 0.0f, tri);
    float label = 0.0f;
    if (best < INFINITY) label = in_box(voxel_box(g, L, vx, vy, vz), along(o, best, d)) ? 1.0f : 0.0f;
    float* in = inputs + (size_t)i * 6;
    in[0] = (o.x - g.bmin[0]) / (g.bmax[0] - g.bmin[0]);
    in[1] = (o.y - g.bmin[1]) / (g.bmax[1] - g.bmin[1]);
    in[2] = (o.z - g.bmin[2]) / (g.bmax[2] - g.bmin[2]);
    in[3] = (d.x + 1.0f) * 0.5f;
    in[4] = (d.y + 1.0f) * 0.5f;
    in[5] = (d.z + 1.0f) * 0.5f;
    labels[i] = label;
}

__device__ __forceinline__ v3 albedo_of(uint64_t albedo_seed, uint32_t tri) {
    uint64_t s = albedo_seed ^ ((uint64_t)tri * kGolden);
    const float a = uniform01(s), b = uniform01(s), c = uniform01(s);
    return {0.1f + 0.8f * a, 0.1f + 0.8f * b, 0.1f + 0.8f * c};
}

// Duff et al. 2017, "Building an Orthonormal Basis, Revisited".
__device__ __forceinline__ void onb(v3 n, v3& b1, v3& b2) {
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    b1 = {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    b2 = {b, sign + n.y * n.y * a, -n.y};
}

__device__ __forceinline__ float ray_eps(v3 o) { return kRayEps * (1.0f + sqrtf(dot(o, o))); }

__global__ __launch_bounds__(kLabelThreads) void throughput_samples_kernel(
    LevelsDev g, const float* __restrict__ tris, uint64_t seed, uint64_t first, uint32_t n, uint32_t spp,
    uint32_t max_depth, uint64_t albedo_seed, float* __restrict__ inputs, float* __restrict__ targets,
    float* __restrict__ valid) {
    const uint32_t i = blockIdx.x * kLabelThreads + threadIdx.x;
    if (i >= n) return;
    uint64_t st = seed ^ ((first + i) * kGolden);
    const uint32_t lod = index_draw(st, g.n_lods);
    const LevelDev& L = g.l[lod];
    const uint32_t slot = index_draw(st, L.n_occupied);
    uint32_t vx, vy, vz;
    voxel_xyz(L.occupied[slot], L.res, vx, vy, vz);
    const v3 wo = sphere_dir(st);
    const v3 wi = sphere_dir(st);
    const Box box = voxel_box(g, L, vx, vy, vz);
    const float U = fmaxf(L.unit[0], fmaxf(L.unit[1], L.unit[2]));
    v3 u_ax, v_ax;
    onb(wo, u_ax, v_ax);
    const v3 d0 = {-wo.x, -wo.y, -wo.z};
    const float back = 2.0f * U;
    v3 sum = {0.0f, 0.0f, 0.0f};
    uint32_t n_valid = 0;
    for (uint32_t j = 0; j < spp; ++j) {
        // entry point: offsets on the plane through centre + 2U wo, normal wo
        bool found = false;
        v3 entry = box.c;
        for (int k = 0; k < kMaxDraws && !found; ++k) {
            const float ox = (uniform01(st) * 2.0f - 1.0f) * U;
            const float oy = (uniform01(st) * 2.0f - 1.0f) * U;
            const v3 o = {((box.c.x + wo.x * back) + u_ax.x * ox) + v_ax.x * oy,
                          ((box.c.y + wo.y * back) + u_ax.y * ox) + v_ax.y * oy,
                          ((box.c.z + wo.z * back) + u_ax.z * ox) + v_ax.z * oy};
            float tn = -INFINITY, tf = INFINITY;
            bool miss = false;
            const float oc[3] = {o.x, o.y, o.z}, dc[3] = {d0.x, d0.y, d0.z};
            const float lo[3] = {box.lo.x, box.lo.y, box.lo.z}, hi[3] = {box.hi.x, box.hi.y, box.hi.z};
            for (int a = 0; a < 3; ++a) {
                if (dc[a] == 0.0f) {
                    if (oc[a] < lo[a] || oc[a] > hi[a]) miss = true;
                    continue;
                }
                float t0 = (lo[a] - oc[a]) / dc[a], t1 = (hi[a] - oc[a]) / dc[a];
                if (t1 < t0) { const float tmp = t0; t0 = t1; t1 = tmp; }
                if (t0 > tn) tn = t0;
                if (t1 < tf) tf = t1;
            }
            if (!miss && tn <= tf && tf >= 0.0f) {
                entry = along(o, tn, d0);
                found = true;
            }
        }
        if (!found) continue;
        v3 o = entry, d = d0, T = {1.0f, 1.0f, 1.0f}, acc = {0.0f, 0.0f, 0.0f};
        float t_min = 0.0f;
        bool path_valid = false;
        for (uint32_t bounce = 0; bounce < max_depth; ++bounce) {
            uint32_t tri = 0;
            const float t = nearest_hit(L, slot, tris, o, d, t_min, tri);
            if (!(t < INFINITY)) break;
            const v3 p = along(o, t, d);
            if (!in_box(box, p)) break;
            if (bounce == 0) path_valid = true;
            const float* tv = tris + (size_t)tri * 9;
            const v3 e1 = {tv[3] - tv[0], tv[4] - tv[1], tv[5] - tv[2]};
            const v3 e2 = {tv[6] - tv[0], tv[7] - tv[1], tv[8] - tv[2]};
            v3 nrm = cross(e1, e2);
            const float inv_len = 1.0f / sqrtf(dot(nrm, nrm));
            nrm = {nrm.x * inv_len, nrm.y * inv_len, nrm.z * inv_len};
            const v3 wout = {-d.x, -d.y, -d.z};
            if (dot(wout, nrm) < 0.0f) nrm = {-nrm.x, -nrm.y, -nrm.z};  // two-sided
            const v3 alb = albedo_of(albedo_seed, tri);
            // next-event estimate along the fixed light direction
            const float eps = ray_eps(p);
            const float cos_i = dot(wi, nrm);
            if (cos_i > 0.0f) {
                uint32_t t2i = 0;
                const float t2 = nearest_hit(L, slot, tris, p, wi, eps, t2i);
                if (!(t2 < INFINITY) || !in_box(box, along(p, t2, wi))) {
                    acc.x = acc.x + (T.x * (alb.x * kInvPi)) * cos_i;
                    acc.y = acc.y + (T.y * (alb.y * kInvPi)) * cos_i;
                    acc.z = acc.z + (T.z * (alb.z * kInvPi)) * cos_i;
                }
            }
            // cosine-weighted continuation: uniform disk point lifted to the hemisphere
            float x = 0.0f, y = 0.0f;
            for (int k = 0; k < kMaxDraws; ++k) {
                const float x1 = uniform01(st) * 2.0f - 1.0f, y1 = uniform01(st) * 2.0f - 1.0f;
                if (x1 * x1 + y1 * y1 <= 1.0f) {
                    x = x1;
                    y = y1;
                    break;
                }
            }
            const float zz = sqrtf(fmaxf(0.0f, 1.0f - (x * x + y * y)));
            v3 b1, b2;
            onb(nrm, b1, b2);
            d = {(b1.x * x + b2.x * y) + nrm.x * zz, (b1.y * x + b2.y * y) + nrm.y * zz,
                 (b1.z * x + b2.z * y) + nrm.z * zz};
            T = {T.x * alb.x, T.y * alb.y, T.z * alb.z};
            o = p;
            t_min = eps;
        }
        if (path_valid) {
            sum = {sum.x + acc.x, sum.y + acc.y, sum.z + acc.z};
            ++n_valid;
        }
    }
    float* in = inputs + (size_t)i * 10;
    in[0] = (box.c.x - g.bmin[0]) / (g.bmax[0] - g.bmin[0]);
    in[1] = (box.c.y - g.bmin[1]) / (g.bmax[1] - g.bmin[1]);
    in[2] = (box.c.z - g.bmin[2]) / (g.bmax[2] - g.bmin[2]);
    in[3] = (wo.x + 1.0f) * 0.5f;
    in[4] = (wo.y + 1.0f) * 0.5f;
    in[5] = (wo.z + 1.0f) * 0.5f;
    in[6] = (wi.x + 1.0f) * 0.5f;
    in[7] = (wi.y + 1.0f) * 0.5f;
    in[8] = (wi.z + 1.0f) * 0.5f;
    in[9] = (float)lod;
    const float nv = (float)n_valid;
    float* out = targets + (size_t)i * 3;
    out[0] = n_valid ? sum.x / nv : 0.0f;
    out[1] = n_valid ? sum.y / nv : 0.0f;
    out[2] = n_valid ? sum.z / nv : 0.0f;
    valid[i] = n_valid ? 1.0f : 0.0f;
}

// glm min(x, y) = (y < x) ? y : x, max(x, y) = (x < y) ? y : x
inline float gmin_h(float x, float y) { return (y < x) ? y : x; }
inline float gmax_h(float x, float y) { return (x < y) ? y : x; }

// Clamped voxel range of a triangle on one axis, as the surface voxelizer
// computes it (cpu_voxelizer.cpp:70-84): (int)clamp((min v - bmin) / unit, 0, G-1).
inline void axis_range(const float* tri, int k, float bmin, float unit, uint32_t G, int& lo, int& hi) {
    const float a = tri[k] - bmin, b = tri[3 + k] - bmin, c = tri[6 + k] - bmin;
    const float gm = (float)(G - 1);
    lo = (int)gmin_h(gmax_h(gmin_h(a, gmin_h(b, c)) / unit, 0.0f), gm);
    hi = (int)gmin_h(gmax_h(gmax_h(a, gmax_h(b, c)) / unit, 0.0f), gm);
}

}  // namespace

struct Labeler {
    struct Level {
        uint32_t res = 0, n_occupied = 0;
        float unit[3] = {0, 0, 0};
        uint64_t n_list = 0;
        DevBuf occupied, list_start, list;
    };
    float bmin[3], bmax[3];
    Level levels[kMaxLods];
    uint32_t n_lods = 0;
    const float* d_tris = nullptr;  // caller-owned, must outlive the labeler

    LevelsDev device_view() const {
        LevelsDev g{};
        for (int k = 0; k < 3; ++k) {
            g.bmin[k] = bmin[k];
            g.bmax[k] = bmax[k];
        }
        g.n_lods = n_lods;
        for (uint32_t l = 0; l < n_lods; ++l) {
            const Level& L = levels[l];
            g.l[l] = LevelDev{L.occupied.as<uint32_t>(), L.list_start.as<uint32_t>(), L.list.as<uint32_t>(), L.res,
                              L.n_occupied, {L.unit[0], L.unit[1], L.unit[2]}};
        }
        return g;
    }
};

Labeler* labeler_create(hipStream_t s, const float* d_tris, uint64_t n_tri, uint32_t res, uint32_t n_lods) {
    NLOD_CHECK(d_tris != nullptr && n_tri > 0, "the mesh needs at least one triangle");
    NLOD_CHECK(res >= 1 && res <= 1024, "label grid resolution must be in [1, 1024]");
    NLOD_CHECK(n_lods >= 1 && n_lods <= (uint32_t)kMaxLods, "n_lods must be in [1, 8]");
    NLOD_CHECK(res % (1u << (n_lods - 1)) == 0, "res must be divisible by 2^(n_lods - 1)");
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
    for (int k = 0; k < 3; ++k) {
        lb->bmin[k] = info.bbox_min[k];
        lb->bmax[k] = info.bbox_max[k];
    }
    lb->d_tris = d_tris;
    lb->n_lods = n_lods;

    // level-0 occupancy as bytes, then OR over 2^l blocks per level
    const uint64_t n_vox0 = (uint64_t)res * res * res;
    std::vector<uint8_t> occ0(n_vox0);
    for (uint64_t v = 0; v < n_vox0; ++v) occ0[v] = (bits[v >> 5] >> (31 - (v & 31))) & 1u;
    for (uint32_t l = 0; l < n_lods; ++l) {
        Labeler::Level& L = lb->levels[l];
        const uint32_t R = res >> l;
        L.res = R;
        for (int k = 0; k < 3; ++k) L.unit[k] = (info.bbox_max[k] - info.bbox_min[k]) / (float)R;
        const uint64_t n_vox = (uint64_t)R * R * R;
        std::vector<uint8_t> occ(n_vox, 0);
        for (uint32_t z = 0; z < res; ++z)
            for (uint32_t y = 0; y < res; ++y)
                for (uint32_t x = 0; x < res; ++x)
                    if (occ0[(uint64_t)x + (uint64_t)y * res + (uint64_t)z * res * res])
                        occ[(uint64_t)(x >> l) + (uint64_t)(y >> l) * R + (uint64_t)(z >> l) * R * R] = 1;
        std::vector<uint32_t> occupied;
        std::vector<int32_t> slot_of(n_vox, -1);
        for (uint64_t v = 0; v < n_vox; ++v)
            if (occ[v]) {
                slot_of[v] = (int32_t)occupied.size();
                occupied.push_back((uint32_t)v);
            }
        NLOD_CHECK(!occupied.empty(), "the mesh occupies no voxel");
        // per occupied voxel: the triangles whose clamped voxel range contains it
        std::vector<uint32_t> count(occupied.size() + 1, 0);
        auto for_voxels = [&](uint64_t t, auto&& f) {
            const float* tri = tris.data() + t * 9;
            for (int k = 0; k < 9; ++k)
                if (!std::isfinite(tri[k])) return;
            int lo[3], hi[3];
            for (int k = 0; k < 3; ++k) axis_range(tri, k, info.bbox_min[k], L.unit[k], R, lo[k], hi[k]);
            for (int z = lo[2]; z <= hi[2]; ++z)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int x = lo[0]; x <= hi[0]; ++x) {
                        const int32_t sl = slot_of[(uint64_t)x + (uint64_t)y * R + (uint64_t)z * R * R];
                        if (sl >= 0) f((uint32_t)sl);
                    }
        };
        for (uint64_t t = 0; t < n_tri; ++t) for_voxels(t, [&](uint32_t sl) { ++count[sl + 1]; });
        for (size_t k = 1; k < count.size(); ++k) count[k] += count[k - 1];
        std::vector<uint32_t> fill(count.begin(), count.end() - 1), lst(count.back());
        for (uint64_t t = 0; t < n_tri; ++t) for_voxels(t, [&](uint32_t sl) { lst[fill[sl]++] = (uint32_t)t; });
        L.n_occupied = (uint32_t)occupied.size();
        L.n_list = lst.size();
        L.occupied.ensure(occupied.size() * sizeof(uint32_t));
        L.list_start.ensure(count.size() * sizeof(uint32_t));
        L.list.ensure(std::max<size_t>(lst.size(), 1) * sizeof(uint32_t));
        NLOD_HIP(hipMemcpyAsync(L.occupied.p, occupied.data(), occupied.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice, s));
        NLOD_HIP(hipMemcpyAsync(L.list_start.p, count.data(), count.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                                s));
        if (!lst.empty())
            NLOD_HIP(hipMemcpyAsync(L.list.p, lst.data(), lst.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        NLOD_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
    }
    return lb.release();
}

void labeler_visibility(const Labeler& lb, hipStream_t s, uint64_t seed, uint64_t first, uint32_t n, float* inputs,
                        float* labels) {
    if (n == 0) return;
    NLOD_CHECK(inputs != nullptr && labels != nullptr, "output pointers are null");
    hipLaunchKernelGGL(visibility_samples_kernel, dim3(div_round_up(n, kLabelThreads)), dim3(kLabelThreads), 0, s,
                       lb.device_view(), lb.d_tris, seed, first, n, inputs, labels);
    NLOD_LAUNCH_CHECK();
}

void labeler_throughput(const Labeler& lb, hipStream_t s, uint64_t seed, uint64_t first, uint32_t n, uint32_t spp,
                        uint32_t max_depth, uint64_t albedo_seed, float* inputs, float* targets, float* valid) {
    if (n == 0) return;
    NLOD_CHECK(inputs != nullptr && targets != nullptr && valid != nullptr, "output pointers are null");
    NLOD_CHECK(spp >= 1 && spp <= 4096, "spp must be in [1, 4096]");
    NLOD_CHECK(max_depth >= 1 && max_depth <= 64, "max_depth must be in [1, 64]");
    hipLaunchKernelGGL(throughput_samples_kernel, dim3(div_round_up(n, kLabelThreads)), dim3(kLabelThreads), 0, s,
                       lb.device_view(), lb.d_tris, seed, first, n, spp, max_depth, albedo_seed, inputs, targets,
                       valid);
    NLOD_LAUNCH_CHECK();
}

}  // namespace nlod

struct nlod_labeler {
    std::unique_ptr<nlod::Labeler> impl;
};

extern "C" {

int nlod_labeler_create(void* stream, const float* d_triangles, uint64_t n_triangles, uint32_t res, uint32_t n_lods,
                        nlod_labeler** out) {
    return nlod::guarded([&] {
        NLOD_CHECK(out != nullptr, "output pointer is null");
        *out = nullptr;
        std::unique_ptr<nlod_labeler> h(new nlod_labeler());
        h->impl.reset(nlod::labeler_create(nlod::as_stream(stream), d_triangles, n_triangles, res, n_lods));
        *out = h.release();
    });
}

int nlod_labeler_destroy(nlod_labeler* labeler) {
    return nlod::guarded([&] { delete labeler; });
}

int nlod_labeler_info(const nlod_labeler* labeler, uint32_t lod, nlod_voxinfo* grid_out, uint32_t* n_occupied_out,
                      uint64_t* n_list_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(labeler != nullptr, "labeler is null");
        const nlod::Labeler& lb = *labeler->impl;
        NLOD_CHECK(lod < lb.n_lods, "lod out of range");
        const nlod::Labeler::Level& L = lb.levels[lod];
        if (grid_out) {
            std::memset(grid_out, 0, sizeof(*grid_out));
            for (int k = 0; k < 3; ++k) {
                grid_out->bbox_min[k] = lb.bmin[k];
                grid_out->bbox_max[k] = lb.bmax[k];
                grid_out->gridsize[k] = L.res;
                grid_out->unit[k] = L.unit[k];
            }
        }
        if (n_occupied_out) *n_occupied_out = L.n_occupied;
        if (n_list_out) *n_list_out = L.n_list;
    });
}

int nlod_labeler_visibility(const nlod_labeler* labeler, void* stream, uint64_t seed, uint64_t first, uint32_t n,
                            float* d_inputs, float* d_labels) {
    return nlod::guarded([&] {
        NLOD_CHECK(labeler != nullptr, "labeler is null");
        nlod::labeler_visibility(*labeler->impl, nlod::as_stream(stream), seed, first, n, d_inputs, d_labels);
    });
}

int nlod_labeler_throughput(const nlod_labeler* labeler, void* stream, uint64_t seed, uint64_t first, uint32_t n,
                            uint32_t spp, uint32_t max_depth, uint64_t albedo_seed, float* d_inputs,
                            float* d_targets, float* d_valid) {
    return nlod::guarded([&] {
        NLOD_CHECK(labeler != nullptr, "labeler is null");
        nlod::labeler_throughput(*labeler->impl, nlod::as_stream(stream), seed, first, n, spp, max_depth, albedo_seed,
                                 d_inputs, d_targets, d_valid);
    });
}

}  // extern "C"
