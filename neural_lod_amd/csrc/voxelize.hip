// Triangle -> bit-packed occupancy voxelization for gfx950.
//
// Surface mode follows the Schwarz-Seidel conservative triangle/box test of
// neural_lod/cuda_voxelizer/src/cpu_voxelizer.cpp:39-193; solid mode follows
// the YZ-projection parity fill of cpu_voxelizer.cpp:252-328.  Results are
// bit-identical to those CPU routines: every floating-point operation below is
// the same IEEE single (or, where the reference mixes in double literals,
// double) operation in the same order, with contraction disabled.
//
// GPU design: a workgroup of 256 lanes streams 256 consecutive triangles
// through LDS with coalesced loads (36 B per triangle, read as one contiguous
// 9 KiB run), then each lane owns one triangle.  Set bits are merged per
// 32-bit word in a register before one atomicOr per word, so a row of x
// candidates that falls in one word costs one atomic instead of one per voxel.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "nlod_common.h"

namespace nlod {
namespace {

constexpr int kVoxThreads = 256;
// cpu_voxelizer.cpp:4: float_error is the double literal 0.000001.
constexpr double kFloatError = 0.000001;

struct VoxParams {
    float bmin[3];
    float unit[3];
    uint32_t G;
    int morton;
    uint64_t n_tri;
};

struct f3 { float x, y, z; };
struct f2 { float x, y; };

__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) {
    float t0 = a.x * b.x, t1 = a.y * b.y, t2 = a.z * b.z;
    return (t0 + t1) + t2;
}
__device__ __forceinline__ float dot2(f2 a, f2 b) {
    float t0 = a.x * b.x, t1 = a.y * b.y;
    return t0 + t1;
}
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ f3 normalize3(f3 v) {
    float inv = 1.0f / sqrtf(dot3(v, v));  // correctly rounded div and sqrt (HIP default)
    return {v.x * inv, v.y * inv, v.z * inv};
}
// min(x,y) = (y < x) ? y : x ; max(x,y) = (x < y) ? y : x
__device__ __forceinline__ float gmin(float x, float y) { return (y < x) ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return (x < y) ? y : x; }
__device__ __forceinline__ float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

__device__ __forceinline__ uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    auto spread = [](uint64_t v) {
        v &= 0x1fffffull;
        v = (v | (v << 32)) & 0x1f00000000ffffull;
        v = (v | (v << 16)) & 0x1f0000ff0000ffull;
        v = (v | (v << 8)) & 0x100f00f00f00f00full;
        v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
        v = (v | (v << 2)) & 0x1249249249249249ull;
        return v;
    };
    return spread(x) | (spread(y) << 1) | (spread(z) << 2);
}

__device__ __forceinline__ uint64_t vox_index(const VoxParams& p, int x, int y, int z) {
    if (p.morton) return morton3((uint32_t)x, (uint32_t)y, (uint32_t)z);
    return (uint64_t)x + (uint64_t)y * p.G + (uint64_t)z * p.G * p.G;
}

// Register-buffered bit writer: merges bits of the same word, one atomic per word.
template <bool XOR>
struct WordWriter {
    uint32_t* table;
    uint64_t word = ~0ull;
    uint32_t mask = 0;
    __device__ void flush() {
        if (mask) {
            if (XOR) atomicXor(table + word, mask);
            else atomicOr(table + word, mask);
        }
        mask = 0;
    }
    __device__ void set(uint64_t index) {
        uint64_t w = index >> 5;
        if (w != word) {
            flush();
            word = w;
        }
        mask |= 1u << (31u - (uint32_t)(index & 31u));
    }
    // XOR (or OR) bits [first, last] of a contiguous index range.
    __device__ void set_range(uint64_t first, uint64_t last) {
        flush();
        uint64_t w0 = first >> 5, w1 = last >> 5;
        for (uint64_t w = w0; w <= w1; ++w) {
            uint32_t lo = (w == w0) ? (uint32_t)(first & 31u) : 0u;
            uint32_t hi = (w == w1) ? (uint32_t)(last & 31u) : 31u;
            // bit b of the word (MSB-first) is index position 31-b; positions lo..hi
            uint32_t m = (0xffffffffu >> lo) & (0xffffffffu << (31u - hi));
            if (XOR) atomicXor(table + w, m);
            else atomicOr(table + w, m);
        }
        word = ~0ull;
    }
};

__device__ __forceinline__ bool tri_finite(const float* t) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 9; ++k) ok = ok && isfinite(t[k]);
    return ok;
}

// cpu_voxelizer.cpp:63-186 for one triangle (vertices t[0..8], untranslated).
__device__ void surface_triangle(const VoxParams& p, const float* t, uint32_t* table) {
    if (!tri_finite(t)) return;
    const f3 unit = {p.unit[0], p.unit[1], p.unit[2]};
    const f3 delta_p = unit;
    const f3 grid_max = {(float)(p.G - 1), (float)(p.G - 1), (float)(p.G - 1)};
    const f3 bmin = {p.bmin[0], p.bmin[1], p.bmin[2]};
    f3 v0 = sub3({t[0], t[1], t[2]}, bmin);
    f3 v1 = sub3({t[3], t[4], t[5]}, bmin);
    f3 v2 = sub3({t[6], t[7], t[8]}, bmin);
    f3 e0 = sub3(v1, v0), e1 = sub3(v2, v1), e2 = sub3(v0, v2);
    f3 n = normalize3(cross3(e0, e1));

    f3 wmin = {gmin(v0.x, gmin(v1.x, v2.x)), gmin(v0.y, gmin(v1.y, v2.y)), gmin(v0.z, gmin(v1.z, v2.z))};
    f3 wmax = {gmax(v0.x, gmax(v1.x, v2.x)), gmax(v0.y, gmax(v1.y, v2.y)), gmax(v0.z, gmax(v1.z, v2.z))};
    const int x0 = (int)gclamp(wmin.x / unit.x, 0.0f, grid_max.x);
    const int y0 = (int)gclamp(wmin.y / unit.y, 0.0f, grid_max.y);
    const int z0 = (int)gclamp(wmin.z / unit.z, 0.0f, grid_max.z);
    const int x1 = (int)gclamp(wmax.x / unit.x, 0.0f, grid_max.x);
    const int y1 = (int)gclamp(wmax.y / unit.y, 0.0f, grid_max.y);
    const int z1 = (int)gclamp(wmax.z / unit.z, 0.0f, grid_max.z);

    f3 c = {0.0f, 0.0f, 0.0f};
    if (n.x > 0.0f) c.x = unit.x;
    if (n.y > 0.0f) c.y = unit.y;
    if (n.z > 0.0f) c.z = unit.z;
    const float d1 = dot3(n, sub3(c, v0));
    const float d2 = dot3(n, sub3(sub3(delta_p, c), v0));

    f2 nxy[3] = {{-1.0f * e0.y, e0.x}, {-1.0f * e1.y, e1.x}, {-1.0f * e2.y, e2.x}};
    if (n.z < 0.0f) for (int k = 0; k < 3; ++k) nxy[k] = {-nxy[k].x, -nxy[k].y};
    f2 nyz[3] = {{-1.0f * e0.z, e0.y}, {-1.0f * e1.z, e1.y}, {-1.0f * e2.z, e2.y}};
    if (n.x < 0.0f) for (int k = 0; k < 3; ++k) nyz[k] = {-nyz[k].x, -nyz[k].y};
    f2 nzx[3] = {{-1.0f * e0.x, e0.z}, {-1.0f * e1.x, e1.z}, {-1.0f * e2.x, e2.z}};
    if (n.y < 0.0f) for (int k = 0; k < 3; ++k) nzx[k] = {-nzx[k].x, -nzx[k].y};
    const f3 vv[3] = {v0, v1, v2};
    float dxy[3], dyz[3], dzx[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        dxy[k] = ((-1.0f * dot2(nxy[k], {vv[k].x, vv[k].y})) + gmax(0.0f, unit.x * nxy[k].x)) + gmax(0.0f, unit.y * nxy[k].y);
        dyz[k] = ((-1.0f * dot2(nyz[k], {vv[k].y, vv[k].z})) + gmax(0.0f, unit.y * nyz[k].x)) + gmax(0.0f, unit.z * nyz[k].y);
        // The CPU path weights the z component by unit.x and the x component
        // by unit.z here (cpu_voxelizer.cpp:133-135); bit-exactness keeps it.
        dzx[k] = ((-1.0f * dot2(nzx[k], {vv[k].z, vv[k].x})) + gmax(0.0f, unit.x * nzx[k].x)) + gmax(0.0f, unit.z * nzx[k].y);
    }

    WordWriter<false> out{table};
    for (int z = z0; z <= z1; ++z) {
        for (int y = y0; y <= y1; ++y) {
            for (int x = x0; x <= x1; ++x) {
                const f3 pt = {(float)x * unit.x, (float)y * unit.y, (float)z * unit.z};
                const float ndp = dot3(n, pt);
                if (((ndp + d1) * (ndp + d2)) > 0.0f) continue;
                const f2 pxy = {pt.x, pt.y}, pyz = {pt.y, pt.z}, pzx = {pt.z, pt.x};
                if ((dot2(nxy[0], pxy) + dxy[0]) < 0.0f) continue;
                if ((dot2(nxy[1], pxy) + dxy[1]) < 0.0f) continue;
                if ((dot2(nxy[2], pxy) + dxy[2]) < 0.0f) continue;
                if ((dot2(nyz[0], pyz) + dyz[0]) < 0.0f) continue;
                if ((dot2(nyz[1], pyz) + dyz[1]) < 0.0f) continue;
                if ((dot2(nyz[2], pyz) + dyz[2]) < 0.0f) continue;
                if ((dot2(nzx[0], pzx) + dzx[0]) < 0.0f) continue;
                if ((dot2(nzx[1], pzx) + dzx[1]) < 0.0f) continue;
                if ((dot2(nzx[2], pzx) + dzx[2]) < 0.0f) continue;
                out.set(vox_index(p, x, y, z));
            }
        }
    }
    out.flush();
}

// check_point_triangle (cpu_voxelizer.cpp:228-249)
__device__ __forceinline__ int check_point_triangle(f2 a, f2 b, f2 c, f2 pt) {
    f2 pa = {pt.x - a.x, pt.y - a.y};
    f2 pb = {pt.x - b.x, pt.y - b.y};
    f2 pc = {pt.x - c.x, pt.y - c.y};
    float t1 = pa.x * pb.y - pa.y * pb.x;
    if ((double)fabsf(t1) < kFloatError && pa.x * pb.x <= 0 && pa.y * pb.y <= 0) return 1;
    float t2 = pb.x * pc.y - pb.y * pc.x;
    if ((double)fabsf(t2) < kFloatError && pb.x * pc.x <= 0 && pb.y * pc.y <= 0) return 2;
    float t3 = pc.x * pa.y - pc.y * pa.x;
    if ((double)fabsf(t3) < kFloatError && pc.x * pa.x <= 0 && pc.y * pa.y <= 0) return 3;
    if (t1 * t2 > 0 && t1 * t3 > 0) return 0;
    return -1;
}

__device__ __forceinline__ bool top_left_edge(f2 a, f2 b) {
    return (b.y < a.y) || (b.y == a.y && a.x > b.x);
}

// cpu_voxelizer.cpp:265-325 for one triangle.
__device__ void solid_triangle(const VoxParams& p, const float* t, uint32_t* table) {
    if (!tri_finite(t)) return;
    const f3 bmin = {p.bmin[0], p.bmin[1], p.bmin[2]};
    f3 v0 = sub3({t[0], t[1], t[2]}, bmin);
    f3 v1 = sub3({t[3], t[4], t[5]}, bmin);
    f3 v2 = sub3({t[6], t[7], t[8]}, bmin);
    f3 e0 = sub3(v1, v0), e1 = sub3(v2, v1);
    f3 n = normalize3(cross3(e0, e1));
    if ((double)fabsf(n.x) < kFloatError) return;
    if (!(isfinite(n.x) && isfinite(n.y) && isfinite(n.z))) return;

    f2 a = {v0.y, v0.z}, b = {v1.y, v1.z}, c = {v2.y, v2.z};
    {
        f2 q0 = {b.x - a.x, b.y - a.y}, q1 = {c.x - a.x, c.y - a.y};
        float r = q0.x * q1.y - q1.x * q0.y;
        if (!(r > 0)) { f2 tmp = b; b = c; c = tmp; }
    }
    const float bmax_x = gmax(a.x, gmax(b.x, c.x)), bmax_y = gmax(a.y, gmax(b.y, c.y));
    const float bmin_x = gmin(a.x, gmin(b.x, c.x)), bmin_y = gmin(a.y, gmin(b.y, c.y));
    const float uy = p.unit[1], uz = p.unit[2], ux = p.unit[0];
    const float gmax_y = (float)floor((double)(bmax_x / uy) - 0.5);
    const float gmax_z = (float)floor((double)(bmax_y / uz) - 0.5);
    const float gmin_y = (float)ceil((double)(bmin_x / uy) - 0.5);
    const float gmin_z = (float)ceil((double)(bmin_y / uz) - 0.5);
    const int G1 = (int)p.G - 1;
    // float -> int conversions of in-range values; clamps keep the walk inside
    // the table (no-ops for meshes inside the box)
    int y0 = (int)fmaxf(gmin_y, -1.0f), y1 = (int)fminf(gmax_y, (float)p.G);
    int z0 = (int)fmaxf(gmin_z, -1.0f), z1 = (int)fminf(gmax_z, (float)p.G);
    y0 = max(y0, 0); z0 = max(z0, 0); y1 = min(y1, G1); z1 = min(z1, G1);

    WordWriter<true> out{table};
    for (int y = y0; y <= y1; ++y) {
        for (int z = z0; z <= z1; ++z) {
            const f2 pt = {(float)(((double)y + 0.5) * (double)uy), (float)(((double)z + 0.5) * (double)uz)};
            const int k = check_point_triangle(a, b, c, pt);
            if ((k == 1 && top_left_edge(a, b)) || (k == 2 && top_left_edge(b, c)) ||
                (k == 3 && top_left_edge(c, a)) || (k == 0)) {
                const float xc = (-(n.y * (pt.x - v0.y) + n.z * (pt.y - v0.z)) / n.x + v0.x);
                const double xq = (double)(xc / ux) - 0.5;
                if (!(xq > -1.0)) continue;
                const int xmax = (xq >= (double)G1) ? G1 : (int)xq;
                if (p.morton) {
                    for (int x = 0; x <= xmax; ++x) out.set(vox_index(p, x, y, z));
                    out.flush();
                } else {
                    const uint64_t row = vox_index(p, 0, y, z);
                    out.set_range(row, row + (uint64_t)xmax);
                }
            }
        }
    }
    out.flush();
}

template <bool SOLID>
__global__ __launch_bounds__(kVoxThreads) void vox_kernel(VoxParams p, const float* __restrict__ tris,
                                                          uint32_t* __restrict__ table) {
    __shared__ float s_tri[kVoxThreads * 9];
    for (uint64_t base = (uint64_t)blockIdx.x * kVoxThreads; base < p.n_tri;
         base += (uint64_t)gridDim.x * kVoxThreads) {
        const uint32_t nt = (uint32_t)min<uint64_t>(kVoxThreads, p.n_tri - base);
        const float* src = tris + base * 9;
        for (uint32_t k = threadIdx.x; k < nt * 9; k += kVoxThreads) s_tri[k] = src[k];
        __syncthreads();
        if (threadIdx.x < nt) {
            float t[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) t[k] = s_tri[threadIdx.x * 9 + k];
            if (SOLID) solid_triangle(p, t, table);
            else surface_triangle(p, t, table);
        }
        __syncthreads();
    }
}

__global__ void fill_zero_kernel(uint32_t* __restrict__ table, uint64_t words, int aligned16) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t n4 = aligned16 ? words / 4 : 0;
    uint4* t4 = reinterpret_cast<uint4*>(table);
    for (uint64_t k = i; k < n4; k += stride) t4[k] = make_uint4(0, 0, 0, 0);
    for (uint64_t k = n4 * 4 + i; k < words; k += stride) table[k] = 0;
}

// Order-preserving float <-> uint mapping for exact atomic min/max.
__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline float ord2f(uint32_t u) {
    uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

__global__ void init_bbox_kernel(uint32_t* acc) {
    if (threadIdx.x < 3) { acc[threadIdx.x] = 0xffffffffu; acc[3 + threadIdx.x] = 0u; }
}

__global__ __launch_bounds__(256) void bbox_kernel(const float* __restrict__ tris, uint64_t n_vals, uint32_t* acc) {
    // n_vals = 9 * n_tri floats; component of value k is k % 3
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    const uint64_t n_vert = n_vals / 3;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n_vert;
         v += (uint64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            float f = tris[v * 3 + d];
            if (isnan(f)) continue;
            uint32_t o = f2ord(f);
            mn[d] = min(mn[d], o);
            mx[d] = max(mx[d], o);
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[d] = min(mn[d], (uint32_t)__shfl_xor((int)mn[d], off));
            mx[d] = max(mx[d], (uint32_t)__shfl_xor((int)mx[d], off));
        }
    }
    if ((threadIdx.x & 63) == 0) {
        for (int d = 0; d < 3; ++d) {
            atomicMin(acc + d, mn[d]);
            atomicMax(acc + 3 + d, mx[d]);
        }
    }
}

int device_count_cached() {
    static int n = [] {
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
        return c;
    }();
    return n;
}

// Per-device 6-word accumulator for the bounding-box reduction, allocated once.
uint32_t* bbox_accumulator() {
    static std::mutex mu;
    static std::vector<uint32_t*> per_dev;
    int dev = 0;
    NLOD_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)per_dev.size() <= dev) per_dev.resize(dev + 1, nullptr);
    if (!per_dev[dev]) NLOD_HIP(hipMalloc(&per_dev[dev], 8 * sizeof(uint32_t)));
    return per_dev[dev];
}

void validate_info(const nlod_voxinfo* info, uint32_t flags) {
    NLOD_CHECK(info != nullptr, "voxinfo is null");
    const uint32_t G = info->gridsize[0];
    NLOD_CHECK(G > 0 && info->gridsize[1] == G && info->gridsize[2] == G,
               "only cubic grids are supported (gridsize.x == gridsize.y == gridsize.z > 0)");
    NLOD_CHECK(G <= 65536, "grid resolution above 65536 is not supported");
    if (flags & NLOD_VOX_MORTON) NLOD_CHECK((G & (G - 1)) == 0, "Morton order needs a power-of-two grid");
    for (int i = 0; i < 3; ++i)
        NLOD_CHECK(std::isfinite(info->unit[i]) && info->unit[i] > 0.0f && std::isfinite(info->bbox_min[i]),
                   "voxinfo has a non-positive or non-finite unit / box");
}

}  // namespace

void voxinfo_from_bbox(const float bmin[3], const float bmax[3], uint32_t G, uint32_t flags, uint64_t n_tri,
                       nlod_voxinfo* info) {
    NLOD_CHECK(G > 0, "grid resolution must be positive");
    float amin[3] = {bmin[0], bmin[1], bmin[2]}, amax[3] = {bmax[0], bmax[1], bmax[2]};
    float len[3];
    for (int i = 0; i < 3; ++i) {
        NLOD_CHECK(std::isfinite(bmin[i]) && std::isfinite(bmax[i]) && bmax[i] >= bmin[i], "invalid bounding box");
        len[i] = bmax[i] - bmin[i];
    }
    // createMeshBBCube (util.h:87-112); GLM max(x,y) = (x < y) ? y : x
    const float maxlen = (len[0] < ((len[1] < len[2]) ? len[2] : len[1])) ? ((len[1] < len[2]) ? len[2] : len[1]) : len[0];
    for (int i = 0; i < 3; ++i) {
        if (maxlen == len[i]) continue;
        const float delta = maxlen - len[i];
        amin[i] = bmin[i] - (delta / 2.0f);
        amax[i] = bmax[i] + (delta / 2.0f);
    }
    if (flags & NLOD_VOX_PAD_EPS) {
        const float eps = 1e-4f;  // main.cpp:45
        float e[3];
        for (int i = 0; i < 3; ++i) e[i] = (amax[i] - amin[i]) * eps;
        for (int i = 0; i < 3; ++i) { amin[i] = amin[i] - e[i]; amax[i] = amax[i] + e[i]; }
    }
    std::memset(info, 0, sizeof(*info));
    for (int i = 0; i < 3; ++i) {
        info->bbox_min[i] = amin[i];
        info->bbox_max[i] = amax[i];
        info->gridsize[i] = G;
        info->unit[i] = (amax[i] - amin[i]) / (float)G;
    }
    info->n_triangles = n_tri;
}

void voxelize(hipStream_t stream, const nlod_voxinfo* info, const float* d_tris, uint32_t* d_table, uint32_t flags) {
    validate_info(info, flags);
    NLOD_CHECK(d_table != nullptr, "voxel table pointer is null");
    NLOD_CHECK(info->n_triangles == 0 || d_tris != nullptr, "triangle pointer is null");
    const uint64_t G = info->gridsize[0];
    const uint64_t words = (G * G * G + 31) / 32;
    if (!(flags & NLOD_VOX_ACCUMULATE)) {
        const uint32_t blocks = (uint32_t)std::min<uint64_t>(div_round_up(words / 4 + 1, 256), 8192);
        hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, stream, d_table, words,
                           (int)(((uintptr_t)d_table & 15u) == 0));
        NLOD_LAUNCH_CHECK();
    }
    if (info->n_triangles == 0) return;
    VoxParams p{};
    for (int i = 0; i < 3; ++i) { p.bmin[i] = info->bbox_min[i]; p.unit[i] = info->unit[i]; }
    p.G = (uint32_t)G;
    p.morton = (flags & NLOD_VOX_MORTON) ? 1 : 0;
    p.n_tri = info->n_triangles;
    const uint64_t blocks = std::min<uint64_t>((p.n_tri + kVoxThreads - 1) / kVoxThreads, 1u << 20);
    if (flags & NLOD_VOX_SOLID)
        hipLaunchKernelGGL(vox_kernel<true>, dim3((uint32_t)blocks), dim3(kVoxThreads), 0, stream, p, d_tris, d_table);
    else
        hipLaunchKernelGGL(vox_kernel<false>, dim3((uint32_t)blocks), dim3(kVoxThreads), 0, stream, p, d_tris, d_table);
    NLOD_LAUNCH_CHECK();
}

void mesh_bbox(hipStream_t stream, const float* d_tris, uint64_t n_tri, float bmin[3], float bmax[3]) {
    NLOD_CHECK(n_tri > 0 && d_tris != nullptr, "mesh must have at least one triangle");
    uint32_t* acc = bbox_accumulator();
    hipLaunchKernelGGL(init_bbox_kernel, dim3(1), dim3(64), 0, stream, acc);
    NLOD_LAUNCH_CHECK();
    const uint64_t n_vert = n_tri * 3;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(div_round_up(n_vert, 256), 2048);
    hipLaunchKernelGGL(bbox_kernel, dim3(blocks), dim3(256), 0, stream, d_tris, n_tri * 9, acc);
    NLOD_LAUNCH_CHECK();
    uint32_t h[6];
    NLOD_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, stream));
    NLOD_HIP(hipStreamSynchronize(stream));
    NLOD_CHECK(h[0] != 0xffffffffu, "mesh has no finite vertex coordinate");
    for (int d = 0; d < 3; ++d) {
        bmin[d] = ord2f(h[d]);
        bmax[d] = ord2f(h[3 + d]);
    }
}

}  // namespace nlod

extern "C" {

int nlod_vox_table_words(uint32_t res, uint64_t* words_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(words_out != nullptr, "output pointer is null");
        NLOD_CHECK(res > 0, "grid resolution must be positive");
        const uint64_t g = res;
        *words_out = (g * g * g + 31) / 32;
    });
}

int nlod_voxinfo_from_bbox(const float bbox_min[3], const float bbox_max[3], uint32_t res, uint32_t flags,
                           uint64_t n_triangles, nlod_voxinfo* info_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(bbox_min && bbox_max && info_out, "null pointer argument");
        nlod::voxinfo_from_bbox(bbox_min, bbox_max, res, flags, n_triangles, info_out);
    });
}

int nlod_mesh_bbox(void* stream, const float* d_triangles, uint64_t n_triangles, float bbox_min_out[3],
                   float bbox_max_out[3]) {
    return nlod::guarded([&] {
        NLOD_CHECK(bbox_min_out && bbox_max_out, "null output pointer");
        nlod::mesh_bbox(nlod::as_stream(stream), d_triangles, n_triangles, bbox_min_out, bbox_max_out);
    });
}

int nlod_voxelize(void* stream, const nlod_voxinfo* info, const float* d_triangles, uint32_t* d_table,
                  uint32_t flags) {
    return nlod::guarded([&] { nlod::voxelize(nlod::as_stream(stream), info, d_triangles, d_table, flags); });
}

int nlod_voxelize_mesh(void* stream, const float* d_triangles, uint64_t n_triangles, uint32_t res, uint32_t flags,
                       uint32_t* d_table, nlod_voxinfo* info_out) {
    return nlod::guarded([&] {
        float bmin[3], bmax[3];
        nlod::mesh_bbox(nlod::as_stream(stream), d_triangles, n_triangles, bmin, bmax);
        nlod_voxinfo info;
        nlod::voxinfo_from_bbox(bmin, bmax, res, flags, n_triangles, &info);
        nlod::voxelize(nlod::as_stream(stream), &info, d_triangles, d_table, flags);
        if (info_out) *info_out = info;
    });
}

}  // extern "C"
