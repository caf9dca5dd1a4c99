// Triangle -> bit-packed occupancy voxelization for gfx950.
//
// Surface mode follows the Schwarz-Seidel conservative triangle/box test of
// neural_lod/cuda_voxelizer/src/cpu_voxelizer.cpp:39-193; solid mode follows
// the YZ-projection parity fill of cpu_voxelizer.cpp:252-328.  Results are
// bit-identical to those CPU routines: every floating-point operation below is
// the same IEEE single (or, where the reference mixes in double literals,
// double) operation in the same order, with contraction disabled.
//
// Two kernel paths give the same bits:
//
// Direct: each lane owns one triangle.  Surface mode sweeps x innermost, so
// the set bits of one row merge into one global atomic per 32-bit word.
// Solid mode XORs each sample row [0, xmax] into the table with one atomic
// per word; at about 21 G/s for random 32-bit global atomics on MI355X that
// is its bound.
//
// Tiled (row-major tables): the grid is cut into tiles of T x T voxel rows
// (full x extent) small enough for LDS, where the bits are set by LDS atomics.
// Four launches:
//   1. count    per (block of triangles, tile): list entries, from the same
//               clamped row range the per-triangle routine walks; the tile
//               range of each triangle is kept for step 3.  The same launch
//               clears the table.
//   2. scan     per tile: exclusive scan over the count blocks, and the total
//   3. scatter  each block scans the tile totals (list starts; block 0 also
//               numbers the chunks) and writes its triangles' indices into
//               their tiles' lists
//   4. tiles    one workgroup per chunk of a tile's list runs the reference's
//               routine restricted to the tile's rows into an LDS copy of the
//               tile, then merges its nonzero words into the table (plain
//               stores when the tile has one chunk and owns whole words,
//               atomics otherwise).  Solid fill on word-aligned rows stores one
//               toggle per sample at xmax and turns the toggles into the row
//               ranges with a suffix-XOR scan per row (XOR is linear, so the
//               chunks of a tile still combine by XOR).  Surface mode spreads
//               the voxel columns of a batch of triangles over the lanes.
//               Triangles touching more than kMaxTiles tiles then run in the
//               same launch with global atomics (every merge is an atomic
//               when there are any).
// Morton-ordered output always takes the direct path.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
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

// Bit sinks: merge the bits of one 32-bit word in a register, then apply them
// to the target with one OR (or XOR for the solid parity fill).
__device__ __forceinline__ uint32_t range_mask(uint32_t lo, uint32_t hi) {
    // bit b of a word (MSB-first) is index position 31-b: positions lo..hi
    return (0xffffffffu >> lo) & (0xffffffffu << (31u - hi));
}

// Target: the whole table in global memory (atomics).
template <bool XOR>
struct GlobalSink {
    uint32_t* table;
    uint64_t word = ~0ull;
    uint32_t mask = 0;
    __device__ void apply(uint64_t w, uint32_t m) {
        if (XOR) atomicXor(table + w, m);
        else atomicOr(table + w, m);
    }
    __device__ void flush() {
        if (mask) apply(word, mask);
        mask = 0;
    }
    __device__ void set(uint64_t index, int) {
        const uint64_t w = index >> 5;
        if (w != word) {
            flush();
            word = w;
        }
        mask |= 1u << (31u - (uint32_t)(index & 31u));
    }
    // bits of the contiguous index range [first, last]
    __device__ void set_range(uint64_t first, uint64_t last, int) {
        flush();
        const uint64_t w0 = first >> 5, w1 = last >> 5;
        for (uint64_t w = w0; w <= w1; ++w)
            apply(w, range_mask(w == w0 ? (uint32_t)(first & 31u) : 0u, w == w1 ? (uint32_t)(last & 31u) : 31u));
        word = ~0ull;
    }
};

// Target: one tile's words in LDS.  Plane zz of the tile (z = z0 + zz) holds
// the words covering voxel rows y0 .. y0 + T - 1 of that plane, starting at
// word first_word(z); the LDS copy of a plane is `wpp` words long.
// DIFF (solid fill of word-aligned rows): set_range(first, last) toggles only
// the bit of `last`; a suffix-XOR scan along each row afterwards turns the
// toggles into the same bits as XOR-ing every range [row start, last].
template <bool XOR, bool DIFF = false>
struct LdsSink {
    uint32_t* lds;
    uint64_t G;
    int z0, y0;
    uint32_t wpp;
    uint64_t word = ~0ull;  // global word index
    uint32_t mask = 0, slot = 0;
    __device__ uint64_t first_word(int z) const { return ((uint64_t)z * G * G + (uint64_t)y0 * G) >> 5; }
    __device__ void apply(uint32_t at, uint32_t m) {
        if (XOR) atomicXor(lds + at, m);
        else atomicOr(lds + at, m);
    }
    __device__ void flush() {
        if (mask) apply(slot, mask);
        mask = 0;
    }
    __device__ void set(uint64_t index, int z) {
        const uint64_t w = index >> 5;
        if (w != word) {
            flush();
            word = w;
            slot = (uint32_t)(z - z0) * wpp + (uint32_t)(w - first_word(z));
        }
        mask |= 1u << (31u - (uint32_t)(index & 31u));
    }
    __device__ void set_range(uint64_t first, uint64_t last, int z) {
        flush();
        const uint32_t base = (uint32_t)(z - z0) * wpp;
        const uint64_t fw = first_word(z);
        if (DIFF) {
            apply(base + (uint32_t)((last >> 5) - fw), 1u << (31u - (uint32_t)(last & 31u)));
            word = ~0ull;
            return;
        }
        const uint64_t w0 = first >> 5, w1 = last >> 5;
        for (uint64_t w = w0; w <= w1; ++w)
            apply(base + (uint32_t)(w - fw),
                  range_mask(w == w0 ? (uint32_t)(first & 31u) : 0u, w == w1 ? (uint32_t)(last & 31u) : 31u));
        word = ~0ull;
    }
};

// Row range (inclusive) a per-triangle routine may touch; empty if y0 > y1.
struct Rows {
    int y0, y1, z0, z1;
};

__device__ __forceinline__ bool tri_finite(const float* t) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 9; ++k) ok = ok && isfinite(t[k]);
    return ok;
}

// Voxel range of a triangle in surface mode (cpu_voxelizer.cpp:70-84).
__device__ Rows surface_rows(const VoxParams& p, const float* t) {
    if (!tri_finite(t)) return {1, 0, 1, 0};
    const f3 unit = {p.unit[0], p.unit[1], p.unit[2]};
    const f3 bmin = {p.bmin[0], p.bmin[1], p.bmin[2]};
    const float gm = (float)(p.G - 1);
    f3 v0 = sub3({t[0], t[1], t[2]}, bmin);
    f3 v1 = sub3({t[3], t[4], t[5]}, bmin);
    f3 v2 = sub3({t[6], t[7], t[8]}, bmin);
    const float ymn = gmin(v0.y, gmin(v1.y, v2.y)), zmn = gmin(v0.z, gmin(v1.z, v2.z));
    const float ymx = gmax(v0.y, gmax(v1.y, v2.y)), zmx = gmax(v0.z, gmax(v1.z, v2.z));
    return {(int)gclamp(ymn / unit.y, 0.0f, gm), (int)gclamp(ymx / unit.y, 0.0f, gm),
            (int)gclamp(zmn / unit.z, 0.0f, gm), (int)gclamp(zmx / unit.z, 0.0f, gm)};
}

// Per-triangle constants of the surface test (cpu_voxelizer.cpp:63-136),
// clipped bounding box and sweep set-up; kept in LDS so that the columns of
// one triangle can run on different lanes.
struct SurfTri {
    float n[3];
    float d1, d2;
    f2 nv[3][3];     // edge normals of the projections xy, yz, zx
    float dv[3][3];  // their offsets
    float lo_ndp, hi_ndp, inv_k;
    uint32_t range[3];  // per axis: lo | hi << 16 (clipped bounding box)
    int axis;           // sweep axis (innermost), -1: no voxel to test
    int slab;           // plane-test interval valid
};

__device__ __forceinline__ int range_lo(uint32_t r) { return (int)(r & 0xffffu); }
__device__ __forceinline__ int range_hi(uint32_t r) { return (int)(r >> 16); }

__device__ __forceinline__ bool edge_tests(const f2* nv, const float* dv, f2 q) {
    // all three edge functions of one projection must be >= 0 (cpu_voxelizer.cpp:164-180)
    return !((dot2(nv[0], q) + dv[0]) < 0.0f) && !((dot2(nv[1], q) + dv[1]) < 0.0f) &&
           !((dot2(nv[2], q) + dv[2]) < 0.0f);
}

// cpu_voxelizer.cpp:63-136 for one triangle (vertices t[0..8], untranslated):
// everything the voxel loop needs, with the loop restricted to rows y in
// [clip.y0, clip.y1] and planes z in [clip.z0, clip.z1].
//
// Sweep: voxels whose plane test provably fails are skipped; every other
// voxel runs the reference's tests unchanged, so the set bits are the same.
// A voxel fails when (ndp + d1) and (ndp + d2) have one strict sign, which
// holds when the exact ndp lies outside [L - M, H + M] (L, H = min, max of
// -d1, -d2).  S bounds |n.x px| + |n.y py| + |n.z pz| over the table and M =
// 1e-5 (S + |d1| + |d2|): the float rounding of ndp is below 4 * 2^-24 * S,
// the float arithmetic that finds the interval (below) errs by less than
// 1e-6 M, and M^2 is a normal float, so the rounded product is then > 0.
// Sweeping along the axis where |n| is largest leaves a short interval per
// column of the other two axes.
//
// x_only: always sweep along x (the reference's loop order; consecutive set
// bits then fall in one word), with the interval only where |n.x| is largest.
__device__ SurfTri surface_setup(const VoxParams& p, const float* t, const Rows clip, bool x_only = false) {
    SurfTri st;
    st.axis = -1;
    if (!tri_finite(t)) return st;
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
    const int y0 = max((int)gclamp(wmin.y / unit.y, 0.0f, grid_max.y), clip.y0);
    const int z0 = max((int)gclamp(wmin.z / unit.z, 0.0f, grid_max.z), clip.z0);
    const int x1 = (int)gclamp(wmax.x / unit.x, 0.0f, grid_max.x);
    const int y1 = min((int)gclamp(wmax.y / unit.y, 0.0f, grid_max.y), clip.y1);
    const int z1 = min((int)gclamp(wmax.z / unit.z, 0.0f, grid_max.z), clip.z1);
    if (y0 > y1 || z0 > z1) return st;

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
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        st.nv[0][k] = nxy[k];
        st.nv[1][k] = nyz[k];
        st.nv[2][k] = nzx[k];
        st.dv[0][k] = ((-1.0f * dot2(nxy[k], {vv[k].x, vv[k].y})) + gmax(0.0f, unit.x * nxy[k].x)) + gmax(0.0f, unit.y * nxy[k].y);
        st.dv[1][k] = ((-1.0f * dot2(nyz[k], {vv[k].y, vv[k].z})) + gmax(0.0f, unit.y * nyz[k].x)) + gmax(0.0f, unit.z * nyz[k].y);
        // The CPU path weights the z component by unit.x and the x component
        // by unit.z here (cpu_voxelizer.cpp:133-135); bit-exactness keeps it.
        st.dv[2][k] = ((-1.0f * dot2(nzx[k], {vv[k].z, vv[k].x})) + gmax(0.0f, unit.x * nzx[k].x)) + gmax(0.0f, unit.z * nzx[k].y);
    }
    st.n[0] = n.x;
    st.n[1] = n.y;
    st.n[2] = n.z;
    st.d1 = d1;
    st.d2 = d2;
    st.range[0] = (uint32_t)x0 | ((uint32_t)x1 << 16);
    st.range[1] = (uint32_t)y0 | ((uint32_t)y1 << 16);
    st.range[2] = (uint32_t)z0 | ((uint32_t)z1 << 16);

    const float S = fabsf(n.x) * ((float)p.G * unit.x) + fabsf(n.y) * ((float)p.G * unit.y) +
                    fabsf(n.z) * ((float)p.G * unit.z);
    const float M = 1e-5f * ((S + fabsf(d1)) + fabsf(d2));
    st.slab = isfinite(M) && S > 1e-12f;
    st.lo_ndp = gmin(-d1, -d2) - M;
    st.hi_ndp = gmax(-d1, -d2) + M;
    int axis = 0;
    if (st.slab) {
        const float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
        axis = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    }
    if (x_only && axis != 0) {
        axis = 0;
        st.slab = 0;
    }
    st.axis = axis;
    const float un[3] = {unit.x, unit.y, unit.z};
    st.inv_k = 1.0f / (st.n[axis] * un[axis]);  // |n[axis]| >= 1/sqrt(3) when slab
    return st;
}

// Number of columns (voxel lines along the sweep axis) of a set-up triangle.
template <int A>
__device__ __forceinline__ uint32_t surface_columns(const SurfTri& st) {
    constexpr int U = A == 2 ? 1 : 2;
    constexpr int V = A == 0 ? 1 : 0;
    if (st.axis != A) return 0;
    return (uint32_t)(range_hi(st.range[U]) - range_lo(st.range[U]) + 1) *
           (uint32_t)(range_hi(st.range[V]) - range_lo(st.range[V]) + 1);
}

// Column j of a set-up triangle with sweep axis A: the projection tests that
// do not involve A run once, the plane test bounds A to an interval, and each
// voxel left runs the plane test and the remaining projection tests as the
// reference loop body (cpu_voxelizer.cpp:148-183).
template <int A, class Sink>
__device__ void surface_column(const VoxParams& p, const SurfTri& st, uint32_t j, Sink& out) {
    constexpr int U = A == 2 ? 1 : 2;  // outer axis
    constexpr int V = A == 0 ? 1 : 0;  // inner column axis
    const float un[3] = {p.unit[0], p.unit[1], p.unit[2]};
    const uint32_t nv_len = (uint32_t)(range_hi(st.range[V]) - range_lo(st.range[V]) + 1);
    const int u = range_lo(st.range[U]) + (int)(j / nv_len);
    const int v = range_lo(st.range[V]) + (int)(j % nv_len);
    const float pu = (float)u * un[U];
    const float pv = (float)v * un[V];
    // the projection onto the plane of U and V: (y, z), (z, x), (x, y)
    if (A == 0 && !edge_tests(st.nv[1], st.dv[1], f2{pv, pu})) return;
    if (A == 1 && !edge_tests(st.nv[2], st.dv[2], f2{pu, pv})) return;
    if (A == 2 && !edge_tests(st.nv[0], st.dv[0], f2{pv, pu})) return;
    int as = range_lo(st.range[A]), ae = range_hi(st.range[A]);
    if (st.slab) {
        const float C = st.n[U] * pu + st.n[V] * pv;
        const float lo_k = st.lo_ndp - C, hi_k = st.hi_ndp - C;  // need n[A] un[A] a in [lo_k, hi_k]
        float xa = lo_k * st.inv_k, xb = hi_k * st.inv_k;
        if (st.inv_k < 0.0f) { const float tmp = xa; xa = xb; xb = tmp; }
        if (xa > (float)ae || xb < (float)as) return;
        if (xa > (float)as) as = max(as, (int)floorf(xa));
        if (xb < (float)ae) ae = min(ae, (int)ceilf(xb));
    }
    const f3 n = {st.n[0], st.n[1], st.n[2]};
    for (int a = as; a <= ae; ++a) {
        int c[3];
        c[A] = a;
        c[U] = u;
        c[V] = v;
        const f3 pt = {(float)c[0] * un[0], (float)c[1] * un[1], (float)c[2] * un[2]};
        const float ndp = dot3(n, pt);
        if (((ndp + st.d1) * (ndp + st.d2)) > 0.0f) continue;
        if (A != 2 && !edge_tests(st.nv[0], st.dv[0], f2{pt.x, pt.y})) continue;
        if (A != 0 && !edge_tests(st.nv[1], st.dv[1], f2{pt.y, pt.z})) continue;
        if (A != 1 && !edge_tests(st.nv[2], st.dv[2], f2{pt.z, pt.x})) continue;
        out.set(vox_index(p, c[0], c[1], c[2]), c[2]);
    }
}

// Exclusive prefix sum over the block (kVoxThreads lanes); returns the total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_part, uint32_t& excl) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_part[wave] = x;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kVoxThreads / 64; ++w) {
        const uint32_t pw = s_part[w];
        if (w < wave) before += pw;
        total += pw;
    }
    excl = before + x - v;
    __syncthreads();  // s_part is reused by the next scan
    return total;
}

// Surface voxelization of up to kVoxThreads set-up triangles held in s_st
// (lane i set up triangle i): the columns of all of them are spread over the
// block, one sweep axis at a time.  s_pre: kVoxThreads + 1 words.
template <class Sink>
__device__ void surface_block(const VoxParams& p, const SurfTri* s_st, uint32_t* s_pre, uint32_t* s_part,
                              Sink& out) {
    const SurfTri& mine = s_st[threadIdx.x];
    auto phase = [&](auto axis_tag) {
        constexpr int A = decltype(axis_tag)::value;
        const uint32_t cols = surface_columns<A>(mine);
        uint32_t excl;
        const uint32_t total = block_exclusive_scan(cols, s_part, excl);
        if (total == 0) return;
        s_pre[threadIdx.x] = excl;
        if (threadIdx.x == 0) s_pre[kVoxThreads] = total;
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < total; k += kVoxThreads) {
            uint32_t lo = 0, hi = kVoxThreads;  // s_pre[lo] <= k < s_pre[hi]
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_pre[mid] <= k) lo = mid;
                else hi = mid;
            }
            surface_column<A>(p, s_st[lo], k - s_pre[lo], out);
        }
        out.flush();
        __syncthreads();  // s_pre is reused by the next phase
    };
    phase(std::integral_constant<int, 0>{});
    phase(std::integral_constant<int, 1>{});
    phase(std::integral_constant<int, 2>{});
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

// Per-triangle part of cpu_voxelizer.cpp:265-325: the triangle relative to the
// box, its unit normal, the YZ projection turned counter-clockwise, and the
// sample rows of its projected bounding box (clamped to the table).
struct SolidTri {
    f2 a, b, c;
    f3 n, v0;
    int y0, y1, z0, z1;
    uint32_t top_left;  // bit 0: edge (a, b), bit 1: (b, c), bit 2: (c, a)
    float pad[3];       // 80 bytes: five 16-byte LDS reads
};

// false: the triangle sets nothing (non-finite, or parallel to the x axis)
__device__ __forceinline__ bool solid_setup(const VoxParams& p, const float* t, SolidTri& st) {
    if (!tri_finite(t)) return false;
    const f3 bmin = {p.bmin[0], p.bmin[1], p.bmin[2]};
    f3 v0 = sub3({t[0], t[1], t[2]}, bmin);
    f3 v1 = sub3({t[3], t[4], t[5]}, bmin);
    f3 v2 = sub3({t[6], t[7], t[8]}, bmin);
    f3 e0 = sub3(v1, v0), e1 = sub3(v2, v1);
    f3 n = normalize3(cross3(e0, e1));
    if ((double)fabsf(n.x) < kFloatError) return false;
    if (!(isfinite(n.x) && isfinite(n.y) && isfinite(n.z))) return false;

    f2 a = {v0.y, v0.z}, b = {v1.y, v1.z}, c = {v2.y, v2.z};
    {
        f2 q0 = {b.x - a.x, b.y - a.y}, q1 = {c.x - a.x, c.y - a.y};
        float r = q0.x * q1.y - q1.x * q0.y;
        if (!(r > 0)) { f2 tmp = b; b = c; c = tmp; }
    }
    const float bmax_x = gmax(a.x, gmax(b.x, c.x)), bmax_y = gmax(a.y, gmax(b.y, c.y));
    const float bmin_x = gmin(a.x, gmin(b.x, c.x)), bmin_y = gmin(a.y, gmin(b.y, c.y));
    const float uy = p.unit[1], uz = p.unit[2];
    const float gmax_y = (float)floor((double)(bmax_x / uy) - 0.5);
    const float gmax_z = (float)floor((double)(bmax_y / uz) - 0.5);
    const float gmin_y = (float)ceil((double)(bmin_x / uy) - 0.5);
    const float gmin_z = (float)ceil((double)(bmin_y / uz) - 0.5);
    const int G1 = (int)p.G - 1;
    // float -> int conversions of in-range values; clamps keep the walk inside
    // the table (no-ops for meshes inside the box)
    int y0 = (int)fmaxf(gmin_y, -1.0f), y1 = (int)fminf(gmax_y, (float)p.G);
    int z0 = (int)fmaxf(gmin_z, -1.0f), z1 = (int)fminf(gmax_z, (float)p.G);
    st.a = a;
    st.b = b;
    st.c = c;
    st.n = n;
    st.v0 = v0;
    st.y0 = max(y0, 0);
    st.z0 = max(z0, 0);
    st.y1 = min(y1, G1);
    st.z1 = min(z1, G1);
    st.top_left = (top_left_edge(a, b) ? 1u : 0u) | (top_left_edge(b, c) ? 2u : 0u) | (top_left_edge(c, a) ? 4u : 0u);
    return true;
}

// One sample row (y, z): the parity test of cpu_voxelizer.cpp:293-320.
template <class Sink>
__device__ __forceinline__ void solid_sample(const VoxParams& p, const SolidTri& st, int y, int z, Sink* out) {
    const float uy = p.unit[1], uz = p.unit[2], ux = p.unit[0];
    const f2 pt = {(float)(((double)y + 0.5) * (double)uy), (float)(((double)z + 0.5) * (double)uz)};
    const int k = check_point_triangle(st.a, st.b, st.c, pt);
    if (!((k == 1 && (st.top_left & 1u)) || (k == 2 && (st.top_left & 2u)) || (k == 3 && (st.top_left & 4u)) ||
          (k == 0)))
        return;
    const f3 n = st.n, v0 = st.v0;
    const float xc = (-(n.y * (pt.x - v0.y) + n.z * (pt.y - v0.z)) / n.x + v0.x);
    const double xq = (double)(xc / ux) - 0.5;
    if (!(xq > -1.0)) return;
    const int G1 = (int)p.G - 1;
    const int xmax = (xq >= (double)G1) ? G1 : (int)xq;
    if (p.morton) {
        for (int x = 0; x <= xmax; ++x) out->set(vox_index(p, x, y, z), z);
        out->flush();
    } else {
        const uint64_t row = vox_index(p, 0, y, z);
        out->set_range(row, row + (uint64_t)xmax, z);
    }
}

// cpu_voxelizer.cpp:265-325 for one triangle, restricted to sample rows y in
// [clip.y0, clip.y1] and z in [clip.z0, clip.z1].  ROWS_ONLY: return the rows
// the walk would visit instead (empty for triangles that set nothing).
template <bool ROWS_ONLY, class Sink>
__device__ Rows solid_triangle(const VoxParams& p, const float* t, Sink* out, const Rows clip) {
    SolidTri st;
    if (!solid_setup(p, t, st)) return {1, 0, 1, 0};
    if (ROWS_ONLY) return {st.y0, st.y1, st.z0, st.z1};
    for (int y = max(st.y0, clip.y0); y <= min(st.y1, clip.y1); ++y)
        for (int z = max(st.z0, clip.z0); z <= min(st.z1, clip.z1); ++z) solid_sample(p, st, y, z, out);
    out->flush();
    return {st.y0, st.y1, st.z0, st.z1};
}

constexpr int kMaxTiles = 9;            // triangles touching more tiles take the direct path
constexpr uint32_t kTileWords = 16384;  // max LDS words of one tile (64 KiB)
constexpr uint32_t kMaxTileCount = 16384;
constexpr uint32_t kMinTile = 16;

// Tile geometry of a row-major table: tiles of T x T rows (full x extent).
struct TileGeo {
    uint32_t T, tiles_y, wpp;  // wpp: LDS words per plane of a tile
    uint32_t n_tiles;
    uint32_t tris_per_block;   // triangles per count / scatter block
    uint32_t chunk;            // list entries per tile workgroup
    int aligned;               // G % 32 == 0: a tile owns whole words, rows are whole words
    int accumulate;            // the table keeps its earlier bits
};

// Per-triangle route, written by the count pass: the tile range packed as
// 4 x 16 bits (ty0, ty1, tz0, tz1), or one of these markers.
constexpr uint64_t kRouteNone = ~0ull;
constexpr uint64_t kRouteDirect = ~0ull - 1;

template <bool SOLID>
__device__ __forceinline__ Rows rows_of(const VoxParams& p, const float* t) {
    if constexpr (SOLID) return solid_triangle<true, GlobalSink<true>>(p, t, nullptr, Rows{0, -1, 0, -1});
    else return surface_rows(p, t);
}

__device__ __forceinline__ uint64_t route_of(const Rows& r, const TileGeo& g) {
    if (r.y0 > r.y1 || r.z0 > r.z1) return kRouteNone;
    const uint32_t ty0 = (uint32_t)r.y0 / g.T, ty1 = (uint32_t)r.y1 / g.T;
    const uint32_t tz0 = (uint32_t)r.z0 / g.T, tz1 = (uint32_t)r.z1 / g.T;
    if ((ty1 - ty0 + 1) * (tz1 - tz0 + 1) > (uint32_t)kMaxTiles) return kRouteDirect;
    return (uint64_t)ty0 | ((uint64_t)ty1 << 16) | ((uint64_t)tz0 << 32) | ((uint64_t)tz1 << 48);
}

// Calls f(tile) for every tile of a packed tile range.
template <class Fn>
__device__ __forceinline__ void for_tiles(uint64_t rt, const TileGeo& g, Fn&& f) {
    const uint32_t ty0 = (uint32_t)(rt & 0xffff), ty1 = (uint32_t)((rt >> 16) & 0xffff);
    const uint32_t tz0 = (uint32_t)((rt >> 32) & 0xffff), tz1 = (uint32_t)(rt >> 48);
    for (uint32_t tz = tz0; tz <= tz1; ++tz)
        for (uint32_t ty = ty0; ty <= ty1; ++ty) f(tz * g.tiles_y + ty);
}

// Coalesced copy of triangles [base, base + nt) (nt <= kVoxThreads) into LDS.
__device__ __forceinline__ void stage_triangles(float* s_tri, const float* __restrict__ tris, uint64_t base,
                                                uint32_t nt) {
    const float* src = tris + base * 9;
    for (uint32_t k = threadIdx.x; k < nt * 9; k += kVoxThreads) s_tri[k] = src[k];
}

// Direct path, solid: one triangle per lane, global atomics.
__global__ __launch_bounds__(kVoxThreads) void vox_solid_kernel(VoxParams p, const float* __restrict__ tris,
                                                                uint32_t* __restrict__ table) {
    __shared__ float s_tri[kVoxThreads * 9];
    const Rows all{0, (int)p.G - 1, 0, (int)p.G - 1};
    for (uint64_t base = (uint64_t)blockIdx.x * kVoxThreads; base < p.n_tri;
         base += (uint64_t)gridDim.x * kVoxThreads) {
        const uint32_t nt = (uint32_t)min<uint64_t>(kVoxThreads, p.n_tri - base);
        stage_triangles(s_tri, tris, base, nt);
        __syncthreads();
        if (threadIdx.x < nt) {
            float t[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) t[k] = s_tri[threadIdx.x * 9 + k];
            GlobalSink<true> out{table};
            solid_triangle<false>(p, t, &out, all);
        }
        __syncthreads();
    }
}

// Direct path, surface: swept along x so that the bits of one row merge into
// one global atomic per word.  Each lane sets up one triangle of a batch and
// the batch's x columns are spread over the block (surface_block), so a wave
// does not wait on its largest triangle.
__global__ __launch_bounds__(kVoxThreads) void vox_surface_kernel(VoxParams p, const float* __restrict__ tris,
                                                                  uint32_t* __restrict__ table) {
    __shared__ float s_tri[kVoxThreads * 9];
    __shared__ SurfTri s_st[kVoxThreads];
    __shared__ uint32_t s_pre[kVoxThreads + 1], s_part[kVoxThreads / 64];
    const Rows all{0, (int)p.G - 1, 0, (int)p.G - 1};
    for (uint64_t base = (uint64_t)blockIdx.x * kVoxThreads; base < p.n_tri;
         base += (uint64_t)gridDim.x * kVoxThreads) {
        const uint32_t nt = (uint32_t)min<uint64_t>(kVoxThreads, p.n_tri - base);
        stage_triangles(s_tri, tris, base, nt);
        __syncthreads();
        SurfTri st;
        st.axis = -1;
        if (threadIdx.x < nt) {
            float t[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) t[k] = s_tri[threadIdx.x * 9 + k];
            st = surface_setup(p, t, all, true);
        }
        s_st[threadIdx.x] = st;
        __syncthreads();
        GlobalSink<false> out{table};
        surface_block(p, s_st, s_pre, s_part, out);
    }
}

// One triangle of the direct route inside the tiled path (it spans more than
// kMaxTiles tiles), with global atomics: the solid walk, or its surface
// columns along x.
template <bool SOLID>
__device__ void direct_triangle(const VoxParams& p, const float* __restrict__ tris, uint64_t i,
                                uint32_t* __restrict__ table) {
    float t[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) t[k] = tris[i * 9 + k];
    const Rows all{0, (int)p.G - 1, 0, (int)p.G - 1};
    if constexpr (SOLID) {
        GlobalSink<true> out{table};
        solid_triangle<false>(p, t, &out, all);
    } else {
        const SurfTri st = surface_setup(p, t, all, true);
        if (st.axis != 0) return;
        GlobalSink<false> out{table};
        const uint32_t cols = surface_columns<0>(st);
        for (uint32_t j = 0; j < cols; ++j) surface_column<0>(p, st, j, out);
        out.flush();
    }
}


// ---- tiled path

// counts[tile][block] = list entries of this block's triangles in the tile;
// routes[i] = triangle i's packed tile range or marker; misc[0] += triangles
// of the direct route.
// LDS: n_tiles counters, then the triangle staging buffer.
// It also clears the table (zero_words > 0: a grid-stride share per block,
// as fill_zero_kernel would).
template <bool SOLID>
__global__ __launch_bounds__(kVoxThreads) void tile_count_kernel(VoxParams p, const float* __restrict__ tris,
                                                                 TileGeo geo, uint32_t n_blocks,
                                                                 uint32_t* __restrict__ counts,
                                                                 uint64_t* __restrict__ routes,
                                                                 uint32_t* __restrict__ misc,
                                                                 uint32_t* __restrict__ table, uint64_t zero_words,
                                                                 int aligned16) {
    extern __shared__ uint32_t cnt[];
    float* s_tri = reinterpret_cast<float*>(cnt + geo.n_tiles);
    if (zero_words) {
        const uint64_t i = (uint64_t)blockIdx.x * kVoxThreads + threadIdx.x, stride = (uint64_t)gridDim.x * kVoxThreads;
        const uint64_t n4 = aligned16 ? zero_words / 4 : 0;
        uint4* t4 = reinterpret_cast<uint4*>(table);
        for (uint64_t k = i; k < n4; k += stride) t4[k] = make_uint4(0, 0, 0, 0);
        for (uint64_t k = n4 * 4 + i; k < zero_words; k += stride) table[k] = 0;
    }
    for (uint32_t t = threadIdx.x; t < geo.n_tiles; t += kVoxThreads) cnt[t] = 0;
    const uint64_t b0 = (uint64_t)blockIdx.x * geo.tris_per_block;
    const uint64_t b1 = min<uint64_t>(p.n_tri, b0 + geo.tris_per_block);
    for (uint64_t base = b0; base < b1; base += kVoxThreads) {
        const uint32_t nt = (uint32_t)min<uint64_t>(kVoxThreads, b1 - base);
        __syncthreads();
        stage_triangles(s_tri, tris, base, nt);
        __syncthreads();
        if (threadIdx.x < nt) {
            float t[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) t[k] = s_tri[threadIdx.x * 9 + k];
            const uint64_t rt = route_of(rows_of<SOLID>(p, t), geo);
            routes[base + threadIdx.x] = rt;
            if (rt < kRouteDirect) for_tiles(rt, geo, [&](uint32_t tile) { atomicAdd(&cnt[tile], 1u); });
            else if (rt == kRouteDirect) atomicAdd(&misc[0], 1u);
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < geo.n_tiles; t += kVoxThreads)
        counts[(size_t)t * n_blocks + blockIdx.x] = cnt[t];
}

// Per tile (one block each): exclusive scan of its row of block counts, in
// place, and the row total.
__global__ __launch_bounds__(kVoxThreads) void tile_scan_rows_kernel(uint32_t n_blocks, uint32_t* __restrict__ counts,
                                                                     uint32_t* __restrict__ totals) {
    __shared__ uint32_t part[kVoxThreads / 64];
    uint32_t* row = counts + (size_t)blockIdx.x * n_blocks;
    const uint32_t per = (n_blocks + kVoxThreads - 1) / kVoxThreads;
    const uint32_t b0 = min(n_blocks, threadIdx.x * per), b1 = min(n_blocks, b0 + per);
    uint32_t s = 0;
    for (uint32_t i = b0; i < b1; ++i) s += row[i];
    uint32_t acc;
    const uint32_t total = block_exclusive_scan(s, part, acc);
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
    for (uint32_t i = b0; i < b1; ++i) {
        const uint32_t v = row[i];
        row[i] = acc;
        acc += v;
    }
}

// Writes the index of every tiled triangle into the lists of its tiles (in
// the same block partition as the count pass).  Each block first scans the
// tile totals itself for the first slot of each tile's list.  Block 0 also
// stores those slots and the chunk numbering for the tile pass (chunk_start
// has n_tiles + 1 entries; the last is the number of chunks; chunk_tile[c] is
// the tile of chunk c), and moves the
// count pass's direct-route triangles to misc[1], clearing misc[0] for the
// next job.
__global__ __launch_bounds__(kVoxThreads) void tile_scatter_kernel(uint64_t n_tri, TileGeo geo, uint32_t n_blocks,
                                                                   const uint64_t* __restrict__ routes,
                                                                   const uint32_t* __restrict__ row_off,
                                                                   const uint32_t* __restrict__ totals,
                                                                   uint32_t* __restrict__ start,
                                                                   uint32_t* __restrict__ chunk_start,
                                                                   uint32_t* __restrict__ chunk_tile,
                                                                   uint32_t* __restrict__ misc,
                                                                   uint32_t* __restrict__ idx) {
    extern __shared__ uint32_t cursor[];
    __shared__ uint32_t s_part[kVoxThreads / 64];
    // exclusive scan of the totals, each thread a contiguous run of tiles
    const uint32_t per = (geo.n_tiles + kVoxThreads - 1) / kVoxThreads;
    const uint32_t t0 = min(geo.n_tiles, threadIdx.x * per), t1 = min(geo.n_tiles, t0 + per);
    uint32_t sum = 0;
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t v = totals[t];
        cursor[t] = v;
        sum += v;
    }
    uint32_t excl;
    block_exclusive_scan(sum, s_part, excl);
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t v = cursor[t];
        cursor[t] = excl;
        excl += v;
    }
    if (blockIdx.x == 0) {
        auto chunks_of = [&](uint32_t t) { return (totals[t] + geo.chunk - 1) / geo.chunk; };
        uint32_t csum = 0;
        for (uint32_t t = t0; t < t1; ++t) csum += chunks_of(t);
        uint32_t cex;
        const uint32_t n_chunks = block_exclusive_scan(csum, s_part, cex);
        for (uint32_t t = t0; t < t1; ++t) {
            start[t] = cursor[t];
            chunk_start[t] = cex;
            const uint32_t nc = chunks_of(t);
            for (uint32_t k = 0; k < nc; ++k) chunk_tile[cex + k] = t;
            cex += nc;
        }
        if (threadIdx.x == 0) {
            chunk_start[geo.n_tiles] = n_chunks;
            misc[1] = misc[0];
            misc[0] = 0;
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < geo.n_tiles; t += kVoxThreads)
        cursor[t] += row_off[(size_t)t * n_blocks + blockIdx.x];
    __syncthreads();
    const uint64_t b0 = (uint64_t)blockIdx.x * geo.tris_per_block;
    const uint64_t b1 = min<uint64_t>(n_tri, b0 + geo.tris_per_block);
    // four routes per lane in flight before the first is used
    constexpr int U = 4;
    for (uint64_t i0 = b0 + threadIdx.x; i0 < b1; i0 += U * kVoxThreads) {
        uint64_t rt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kVoxThreads;
            rt[u] = i < b1 ? routes[i] : kRouteNone;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = (uint32_t)(i0 + (uint64_t)u * kVoxThreads);
            if (rt[u] < kRouteDirect)
                for_tiles(rt[u], geo, [&](uint32_t tile) { idx[atomicAdd(&cursor[tile], 1u)] = i; });
        }
    }
}

// Suffix-XOR along each word-aligned row of the LDS tile (solid DIFF form):
// bit x of a row becomes the XOR of the toggles at x' >= x.  Bit 31-b of a
// word is x = 32w + b, so within a word the suffix runs from bit 0 upwards.
__device__ __forceinline__ void rows_suffix_xor(uint32_t* lds, uint32_t n_rows, uint32_t row_words) {
    for (uint32_t r = threadIdx.x; r < n_rows; r += kVoxThreads) {
        uint32_t* row = lds + r * row_words;
        uint32_t carry = 0;
        for (int w = (int)row_words - 1; w >= 0; --w) {
            uint32_t v = row[w];
            v ^= v << 1;
            v ^= v << 2;
            v ^= v << 4;
            v ^= v << 8;
            v ^= v << 16;
            if (carry) v = ~v;
            carry = v >> 31;
            row[w] = v;
        }
    }
}

// Persistent: each iteration takes one chunk of one tile's list, runs the
// per-triangle routine clipped to the tile's rows into an LDS copy of the
// tile, then merges the nonzero words into the table.  Then the triangles of
// the direct route (misc[1] of them), with global atomics; while there are
// any, every merge is an atomic.
template <bool SOLID, bool DIFF>
__global__ __launch_bounds__(kVoxThreads) void tile_vox_kernel(VoxParams p, const float* __restrict__ tris,
                                                               TileGeo geo, const uint32_t* __restrict__ totals,
                                                               const uint32_t* __restrict__ start,
                                                               const uint32_t* __restrict__ chunk_start,
                                                               const uint32_t* __restrict__ chunk_tile,
                                                               const uint32_t* __restrict__ idx,
                                                               const uint64_t* __restrict__ routes,
                                                               const uint32_t* __restrict__ misc,
                                                               uint32_t* __restrict__ table) {
    extern __shared__ uint32_t lds[];
    const uint32_t n_chunks = chunk_start[geo.n_tiles];
    const bool any_direct = misc[1] != 0;
    const uint32_t tile_words = geo.T * geo.wpp;
    const uint64_t GG = (uint64_t)p.G * p.G;
    for (uint32_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        const uint32_t tile = chunk_tile[c];
        const int ty = (int)(tile % geo.tiles_y), tz = (int)(tile / geo.tiles_y);
        const Rows clip{ty * (int)geo.T, min((ty + 1) * (int)geo.T, (int)p.G) - 1, tz * (int)geo.T,
                        min((tz + 1) * (int)geo.T, (int)p.G) - 1};
        const uint32_t begin = start[tile] + (c - chunk_start[tile]) * geo.chunk;
        const uint32_t end = min(start[tile] + totals[tile], begin + geo.chunk);
        // each lane's triangle of the batch at kb; the next batch's is loaded
        // while the current one runs (the first while the tile is cleared)
        auto load_tri = [&](uint32_t kb, float(&t)[9]) {
            const uint32_t k = kb + threadIdx.x;
            if (k >= end) return false;
            const uint64_t i = idx[k];
#pragma unroll
            for (int j = 0; j < 9; ++j) t[j] = tris[i * 9 + j];
            return true;
        };
        float tn[9];
        bool hn = load_tri(begin, tn);
        for (uint32_t k = threadIdx.x; k < tile_words; k += kVoxThreads) lds[k] = 0;
        __syncthreads();
        if constexpr (SOLID) {
            // A batch of kVoxThreads triangles, 64 per wave: each lane sets up one
            // and counts its sample rows inside the tile; a wave scan numbers the
            // wave's samples, and its lanes then take samples in turn from any of
            // its triangles, so a wave runs its mean sample count, not its largest
            // triangle's, and no block barrier separates the batches.  After the
            // tile: set-up triangles (64 per wave), then 65 sample offsets per wave.
            const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            SolidTri* s_st_all = reinterpret_cast<SolidTri*>(lds + tile_words);
            SolidTri* s_st = s_st_all + wave * 64;
            uint32_t* s_off = reinterpret_cast<uint32_t*>(s_st_all + kVoxThreads) + wave * 65;
            LdsSink<true, DIFF> out{lds, (uint64_t)p.G, clip.z0, clip.y0, geo.wpp};
            for (uint32_t kb = begin; kb < end; kb += kVoxThreads) {
                float t[9];
#pragma unroll
                for (int j = 0; j < 9; ++j) t[j] = tn[j];
                const bool have = hn;
                hn = load_tri(kb + kVoxThreads, tn);
                uint32_t cnt = 0;
                if (have) {
                    SolidTri st;
                    if (solid_setup(p, t, st)) {
                        st.y0 = max(st.y0, clip.y0);
                        st.y1 = min(st.y1, clip.y1);
                        st.z0 = max(st.z0, clip.z0);
                        st.z1 = min(st.z1, clip.z1);
                        if (st.y0 <= st.y1 && st.z0 <= st.z1) {
                            cnt = (uint32_t)(st.y1 - st.y0 + 1) * (uint32_t)(st.z1 - st.z0 + 1);
                            s_st[lane] = st;
                        }
                    }
                }
                uint32_t x = cnt;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const uint32_t y = (uint32_t)__shfl_up((int)x, off, 64);
                    if (lane >= (uint32_t)off) x += y;
                }
                const uint32_t total = (uint32_t)__shfl((int)x, 63, 64);
                s_off[lane] = x - cnt;
                if (lane == 63) s_off[64] = total;
                // the wave's own LDS writes, read back by other lanes of the wave
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (uint32_t si = lane; si < total; si += 64) {
                    uint32_t lo = 0, hi = 64;  // s_off[lo] <= si < s_off[hi]
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (s_off[mid] <= si) lo = mid;
                        else hi = mid;
                    }
                    const SolidTri st = s_st[lo];
                    const uint32_t local = si - s_off[lo];
                    const uint32_t nz = (uint32_t)(st.z1 - st.z0 + 1);
                    // local / nz: local < 64 nz and nz <= 64 (tiles are at most 64 rows
                    // a side), where a float reciprocal within 2 ulp floors exactly
                    const uint32_t yi = (uint32_t)(((float)local + 0.5f) * __frcp_rn((float)nz));
                    solid_sample(p, st, st.y0 + (int)yi, st.z0 + (int)(local - yi * nz), &out);
                }
                // the wave's set-ups and offsets are reused by its next batch
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        } else {
            // after the tile: set-up triangles, column prefix, scan partials
            SurfTri* s_st = reinterpret_cast<SurfTri*>(lds + tile_words);
            uint32_t* s_pre = reinterpret_cast<uint32_t*>(s_st + kVoxThreads);
            uint32_t* s_part = s_pre + kVoxThreads + 1;
            for (uint32_t kb = begin; kb < end; kb += kVoxThreads) {
                float t[9];
#pragma unroll
                for (int j = 0; j < 9; ++j) t[j] = tn[j];
                const bool have = hn;
                hn = load_tri(kb + kVoxThreads, tn);
                SurfTri st;
                st.axis = -1;
                if (have) st = surface_setup(p, t, clip);
                s_st[threadIdx.x] = st;
                __syncthreads();
                LdsSink<false> out{lds, (uint64_t)p.G, clip.z0, clip.y0, geo.wpp};
                surface_block(p, s_st, s_pre, s_part, out);
                __syncthreads();
            }
        }
        __syncthreads();
        if constexpr (DIFF) {
            // rows of plane zz start at zz * wpp, each G / 32 words (aligned G)
            rows_suffix_xor(lds, geo.T * geo.T, p.G / 32);
            __syncthreads();
        }
        // a tile with one chunk that owns whole words is the only writer of them,
        // unless triangles of the direct route write into the table as well
        const bool exclusive = geo.aligned && !any_direct && chunk_start[tile + 1] - chunk_start[tile] == 1;
        const uint32_t planes = (uint32_t)(clip.z1 - clip.z0 + 1);
        for (uint32_t k = threadIdx.x; k < planes * geo.wpp; k += kVoxThreads) {
            const uint32_t v = lds[k];
            if (!v) continue;
            const uint32_t zz = k / geo.wpp;
            const uint64_t w = (((uint64_t)(clip.z0 + (int)zz) * GG + (uint64_t)clip.y0 * p.G) >> 5) + (k - zz * geo.wpp);
            if (exclusive) table[w] = geo.accumulate ? (SOLID ? (table[w] ^ v) : (table[w] | v)) : v;
            else if (SOLID) atomicXor(table + w, v);
            else atomicOr(table + w, v);
        }
        __syncthreads();  // LDS is reused by the next chunk
    }
    if (any_direct) {
        for (uint64_t i = (uint64_t)blockIdx.x * kVoxThreads + threadIdx.x; i < p.n_tri;
             i += (uint64_t)gridDim.x * kVoxThreads)
            if (routes[i] == kRouteDirect) direct_triangle<SOLID>(p, tris, i, table);
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

// Bounding box of n_vals = 9 n_tri floats (the component of float k is k % 3).
// With a 16-byte aligned array the block reads tiles of 3 x 256 float4s (3072
// floats, so a tile starts at component 0), each load instruction one
// contiguous run across the block: thread t's float4 t + 256 j holds floats
// whose components are (t + j + e) % 3 for element e, so the thread keeps its
// minima per phase (j + e) % 3 and maps phases to components at the end.  The
// rest, or an unaligned array, goes float by float.  NaN coordinates are skipped.
constexpr int kBboxThreads = 256;

__device__ __forceinline__ void bbox_add(uint32_t* mn, uint32_t* mx, int d, float f) {
    if (isnan(f)) return;
    const uint32_t o = f2ord(f);
    mn[d] = min(mn[d], o);
    mx[d] = max(mx[d], o);
}

__global__ __launch_bounds__(kBboxThreads) void bbox_kernel(const float* __restrict__ tris, uint64_t n_vals,
                                                            bool aligned16, uint32_t* acc) {
    constexpr uint64_t kTile = 3 * kBboxThreads;  // float4s per tile
    uint32_t pmn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, pmx[3] = {0u, 0u, 0u};
    const uint32_t t = threadIdx.x;
    const uint64_t n_tiles = aligned16 ? n_vals / (4 * kTile) : 0;
    const float4* t4 = reinterpret_cast<const float4*>(tris);
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        float4 v[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) v[j] = t4[tile * kTile + t + (uint64_t)kBboxThreads * j];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            bbox_add(pmn, pmx, (j + 0) % 3, v[j].x);
            bbox_add(pmn, pmx, (j + 1) % 3, v[j].y);
            bbox_add(pmn, pmx, (j + 2) % 3, v[j].z);
            bbox_add(pmn, pmx, (j + 3) % 3, v[j].w);
        }
    }
    // phase p holds component (t + p) % 3
    uint32_t mn[3], mx[3];
    const uint32_t r = t % 3;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const uint32_t p = (uint32_t)(d + 3 - (int)r) % 3;
        mn[d] = p == 0 ? pmn[0] : p == 1 ? pmn[1] : pmn[2];
        mx[d] = p == 0 ? pmx[0] : p == 1 ? pmx[1] : pmx[2];
    }
    const uint64_t gid = (uint64_t)blockIdx.x * kBboxThreads + t;
    const uint64_t stride = (uint64_t)gridDim.x * kBboxThreads;
    for (uint64_t k = n_tiles * 4 * kTile + gid; k < n_vals; k += stride) bbox_add(mn, mx, (int)(k % 3), tris[k]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[d] = min(mn[d], (uint32_t)__shfl_xor((int)mn[d], off));
            mx[d] = max(mx[d], (uint32_t)__shfl_xor((int)mx[d], off));
        }
    }
    __shared__ uint32_t wave_mm[kBboxThreads / 64][6];
    if ((t & 63) == 0)
        for (int d = 0; d < 3; ++d) {
            wave_mm[t >> 6][d] = mn[d];
            wave_mm[t >> 6][3 + d] = mx[d];
        }
    __syncthreads();
    if (t < 6) {
        uint32_t m = wave_mm[0][t];
        for (int w = 1; w < kBboxThreads / 64; ++w) m = t < 3 ? min(m, wave_mm[w][t]) : max(m, wave_mm[w][t]);
        if (t < 3) atomicMin(acc + t, m);
        else atomicMax(acc + t, m);
    }
}

// 6-word accumulator of the bounding-box reduction, one per (device, stream),
// allocated on first use: calls on different streams (from different host
// threads) never share one, and calls on one stream are ordered by it.
uint32_t* bbox_accumulator(hipStream_t s) {
    static std::mutex mu;
    static std::map<std::pair<int, hipStream_t>, std::unique_ptr<DevBuf>> per;
    int dev = 0;
    NLOD_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto& b = per[std::make_pair(dev, s)];
    if (!b) {
        b.reset(new DevBuf());
        b->ensure(8 * sizeof(uint32_t));
    }
    return b->as<uint32_t>();
}

// Tile shape for a G^3 row-major table: T x T rows with T the smallest
// size >= kMinTile that keeps the tile count within kMaxTileCount (the count
// pass keeps one LDS counter per tile).  False when such a tile does not fit
// kTileWords LDS words, or (unless forced) for small jobs, where the extra
// passes cost more than the global atomics they save.
bool tile_geometry(uint32_t G, uint64_t n_tri, bool force, bool accumulate, TileGeo& g) {
    if ((!force && n_tri < 4096) || n_tri * kMaxTiles >= (1ull << 32)) return false;
    const bool aligned = (G % 32) == 0;
    uint64_t T = std::min<uint64_t>(G, kMinTile);
    while ((G + T - 1) / T * ((G + T - 1) / T) > kMaxTileCount) ++T;
    const uint64_t wpp = aligned ? T * G / 32 : (T * G + 31) / 32 + 1;
    if (T * wpp > kTileWords || T > 64) return false;  // (T <= 22 follows from the word limit)
    const uint64_t tiles_y = (G + T - 1) / T;
    g.T = (uint32_t)T;
    g.tiles_y = (uint32_t)tiles_y;
    g.wpp = (uint32_t)wpp;
    g.n_tiles = (uint32_t)(tiles_y * tiles_y);
    // count / scatter blocks: 1024 to 4096 as the count matrix (one row of block
    // counts per tile) stays within 8M entries, each a multiple of the staging
    // size; small meshes still get one block per 256 triangles
    const uint64_t target = std::min<uint64_t>(4096, std::max<uint64_t>(1024, (8ull << 20) / g.n_tiles));
    const uint64_t per = std::max<uint64_t>(kVoxThreads, (n_tri + target - 1) / target);
    g.tris_per_block = (uint32_t)((per + kVoxThreads - 1) / kVoxThreads * kVoxThreads);
    // list entries per workgroup: about 16 chunks per CU for ~2 entries per triangle
    const uint64_t chunk = 2 * n_tri / (16ull * (uint64_t)cu_count());
    g.chunk = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(chunk, kVoxThreads), 1u << 16);
    g.aligned = aligned ? 1 : 0;
    g.accumulate = accumulate ? 1 : 0;
    return true;
}

// The tiled path's scratch starts with kMiscWords job counters (misc[0]:
// direct-route triangles, counted by the count pass and cleared by the
// scatter pass, so it is zero between jobs; misc[1]: the same count handed to
// the tile pass), at the same place for every job geometry.
constexpr size_t kMiscWords = 16;

// Tiled-path scratch, one grow-only buffer per (device, stream); a new
// buffer's job counters are cleared on the stream before its first job.
uint8_t* tile_workspace(hipStream_t s, size_t bytes) {
    static std::mutex mu;
    static std::map<std::pair<int, hipStream_t>, std::unique_ptr<DevBuf>> per;
    int dev = 0;
    NLOD_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto& b = per[std::make_pair(dev, s)];
    if (!b) b.reset(new DevBuf());
    if (b->bytes < bytes) {
        NLOD_HIP(hipStreamSynchronize(s));  // earlier calls on this stream may still read it
        b->ensure(bytes);
        hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(64), 0, s, b->as<uint32_t>(), (uint64_t)kMiscWords, 1);
        NLOD_LAUNCH_CHECK();
    }
    return b->as<uint8_t>();
}

// The direct kernels loop over 256-triangle batches.  While the batches fit
// in four rounds of two blocks per CU, the grid stays at two blocks per CU:
// three resident blocks per CU (their occupancy) would finish a batch count
// just above one round with a tail of single blocks (200k triangles: 782
// batches against 768 slots).
template <bool SOLID>
void launch_direct(hipStream_t stream, const VoxParams& p, const float* d_tris, uint32_t* d_table) {
    const uint64_t batches = (p.n_tri + kVoxThreads - 1) / kVoxThreads;
    const uint64_t pair_grid = 2ull * (uint64_t)cu_count();
    const uint64_t blocks = std::min<uint64_t>(batches <= 4 * pair_grid ? std::min(batches, pair_grid) : batches,
                                               1u << 20);
    if (SOLID)
        hipLaunchKernelGGL(vox_solid_kernel, dim3((uint32_t)blocks), dim3(kVoxThreads), 0, stream, p, d_tris, d_table);
    else
        hipLaunchKernelGGL(vox_surface_kernel, dim3((uint32_t)blocks), dim3(kVoxThreads), 0, stream, p, d_tris,
                           d_table);
    NLOD_LAUNCH_CHECK();
}

// Most chunks a job can have: one partial chunk per tile, plus the full ones.
size_t max_chunks(const VoxParams& p, const TileGeo& geo) {
    return (size_t)geo.n_tiles + (size_t)p.n_tri * kMaxTiles / geo.chunk + 1;
}

// Bytes of tiled-path scratch for a job of this geometry: misc | routes (u64
// per triangle) | counts | totals | start | chunk_start (n_tiles + 1) |
// chunk_tile (max_chunks) | idx.
size_t tile_scratch_bytes(const VoxParams& p, const TileGeo& geo) {
    const uint32_t n_blocks = (uint32_t)((p.n_tri + geo.tris_per_block - 1) / geo.tris_per_block);
    const size_t n_counts = (size_t)geo.n_tiles * n_blocks;
    const size_t n_idx = (size_t)p.n_tri * kMaxTiles;
    return (kMiscWords + 2 * (size_t)p.n_tri + n_counts + 3 * (size_t)geo.n_tiles + 1 + max_chunks(p, geo) + n_idx) *
           sizeof(uint32_t);
}

// The tiled path's scratch comes from `workspace(bytes)`: the library's
// per-stream buffer, or the caller-free static array of nlod_voxelize_noalloc
// (zero at load).  Every word of it that a kernel reads was written earlier in
// the same job, except misc[0], which jobs leave zero.
template <bool SOLID, typename Workspace>
void launch_voxelize(hipStream_t stream, const VoxParams& p, const TileGeo& geo, bool tiled, const float* d_tris,
                     uint32_t* d_table, Workspace&& workspace) {
    if (!tiled) {
        launch_direct<SOLID>(stream, p, d_tris, d_table);
        return;
    }
    const uint32_t n_blocks = (uint32_t)((p.n_tri + geo.tris_per_block - 1) / geo.tris_per_block);
    const size_t n_counts = (size_t)geo.n_tiles * n_blocks;
    uint8_t* ws = workspace(tile_scratch_bytes(p, geo));
    uint32_t* misc = reinterpret_cast<uint32_t*>(ws);
    uint64_t* routes = reinterpret_cast<uint64_t*>(misc + kMiscWords);
    uint32_t* counts = reinterpret_cast<uint32_t*>(routes + p.n_tri);
    uint32_t* totals = counts + n_counts;
    uint32_t* start = totals + geo.n_tiles;
    uint32_t* chunk_start = start + geo.n_tiles;
    uint32_t* chunk_tile = chunk_start + geo.n_tiles + 1;
    uint32_t* idx = chunk_tile + max_chunks(p, geo);
    const size_t count_smem = (size_t)geo.n_tiles * 4 + kVoxThreads * 9 * sizeof(float);
    const size_t scatter_smem = (size_t)geo.n_tiles * 4;
    (void)occupancy_blocks(reinterpret_cast<const void*>(tile_count_kernel<SOLID>), kVoxThreads, count_smem);
    (void)occupancy_blocks(reinterpret_cast<const void*>(tile_scatter_kernel), kVoxThreads, scatter_smem);
    const uint64_t zero_words = geo.accumulate ? 0 : ((uint64_t)p.G * p.G * p.G + 31) / 32;
    hipLaunchKernelGGL(tile_count_kernel<SOLID>, dim3(n_blocks), dim3(kVoxThreads), count_smem, stream, p, d_tris,
                       geo, n_blocks, counts, routes, misc, d_table, zero_words,
                       (int)(((uintptr_t)d_table & 15u) == 0));
    NLOD_LAUNCH_CHECK();
    hipLaunchKernelGGL(tile_scan_rows_kernel, dim3(geo.n_tiles), dim3(kVoxThreads), 0, stream, n_blocks, counts,
                       totals);
    NLOD_LAUNCH_CHECK();
    hipLaunchKernelGGL(tile_scatter_kernel, dim3(n_blocks), dim3(kVoxThreads), scatter_smem, stream, p.n_tri, geo,
                       n_blocks, routes, counts, totals, start, chunk_start, chunk_tile, misc, idx);
    NLOD_LAUNCH_CHECK();
    size_t tile_smem = (size_t)geo.T * geo.wpp * sizeof(uint32_t) +
                       (kVoxThreads + 1 + kVoxThreads / 64) * sizeof(uint32_t);
    tile_smem += kVoxThreads * (SOLID ? sizeof(SolidTri) : sizeof(SurfTri));
    auto run_tiles = [&](auto kernel) {
        const int per_cu = occupancy_blocks(reinterpret_cast<const void*>(kernel), kVoxThreads, tile_smem);
        const uint32_t grid = (uint32_t)(cu_count() * per_cu);
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kVoxThreads), tile_smem, stream, p, d_tris, geo, totals, start,
                           chunk_start, chunk_tile, idx, routes, misc, d_table);
        NLOD_LAUNCH_CHECK();
    };
    if (SOLID && geo.aligned) run_tiles(tile_vox_kernel<SOLID, true>);
    else run_tiles(tile_vox_kernel<SOLID, false>);
}

void validate_info(const nlod_voxinfo* info, uint32_t flags) {
    NLOD_CHECK(info != nullptr, "voxinfo is null");
    const uint32_t G = info->gridsize[0];
    NLOD_CHECK(G > 0 && info->gridsize[1] == G && info->gridsize[2] == G,
               "only cubic grids are supported (gridsize.x == gridsize.y == gridsize.z > 0)");
    NLOD_CHECK(G <= 65536, "grid resolution above 65536 is not supported");
    if (flags & NLOD_VOX_MORTON) NLOD_CHECK((G & (G - 1)) == 0, "Morton order needs a power-of-two grid");
    NLOD_CHECK((flags & (NLOD_VOX_DIRECT | NLOD_VOX_TILED)) != (NLOD_VOX_DIRECT | NLOD_VOX_TILED),
               "NLOD_VOX_DIRECT and NLOD_VOX_TILED exclude each other");
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

// File-scope scratch of nlod_voxelize_noalloc: one copy per device, held until
// the process ends.  Host code never passes `scratch` itself to a kernel (in
// host code the name is a stand-in, not a device address); it uses the address
// hipGetSymbolAddress returns on each device.
constexpr size_t kNoAllocScratchBytes = 32u << 20;
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
uint8_t* noalloc_scratch() {
    static std::mutex mu;
    static std::vector<uint8_t*> per_dev;
    int dev = 0;
    NLOD_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)per_dev.size() <= dev) per_dev.resize(dev + 1, nullptr);
    if (!per_dev[dev]) {
        void* ptr = nullptr;
        NLOD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        NLOD_CHECK(ptr != nullptr, "hipGetSymbolAddress returned no address for the voxelizer scratch");
        per_dev[dev] = static_cast<uint8_t*>(ptr);
    }
    return per_dev[dev];
}

void voxelize(hipStream_t stream, const nlod_voxinfo* info, const float* d_tris, uint32_t* d_table, uint32_t flags,
              bool no_alloc) {
    validate_info(info, flags);
    NLOD_CHECK(d_table != nullptr, "voxel table pointer is null");
    NLOD_CHECK(info->n_triangles == 0 || d_tris != nullptr, "triangle pointer is null");
    const uint64_t G = info->gridsize[0];
    const uint64_t words = (G * G * G + 31) / 32;
    VoxParams p{};
    for (int i = 0; i < 3; ++i) { p.bmin[i] = info->bbox_min[i]; p.unit[i] = info->unit[i]; }
    p.G = (uint32_t)G;
    p.morton = (flags & NLOD_VOX_MORTON) ? 1 : 0;
    p.n_tri = info->n_triangles;
    TileGeo geo{};
    // default: tiled for the solid fill (its row ranges make global atomics
    // the bound of the direct kernel); the surface test is compute-bound
    // either way and the direct kernel has no binning passes
    const bool want = (flags & NLOD_VOX_TILED) || ((flags & NLOD_VOX_SOLID) && !(flags & NLOD_VOX_DIRECT));
    const bool tiled = p.n_tri > 0 && !p.morton && want &&
                       tile_geometry(p.G, p.n_tri, (flags & NLOD_VOX_TILED) != 0, (flags & NLOD_VOX_ACCUMULATE) != 0,
                                     geo);
    uint8_t* fixed = nullptr;
    if (no_alloc) {  // checked before any launch: a rejected call leaves the table untouched
        if (tiled && tile_scratch_bytes(p, geo) > kNoAllocScratchBytes)
            fail(NLOD_ERR_INVALID_ARGUMENT,
                 "nlod_voxelize_noalloc: this job needs " + std::to_string(tile_scratch_bytes(p, geo)) +
                     " bytes of tiled-path scratch, more than the " + std::to_string(kNoAllocScratchBytes) +
                     " bytes it holds (pass NLOD_VOX_DIRECT, which needs none, or use nlod_voxelize)");
        fixed = noalloc_scratch();
    }
    if (!(flags & NLOD_VOX_ACCUMULATE) && !tiled) {  // the tiled path's count pass clears the table itself
        const uint32_t blocks = (uint32_t)std::min<uint64_t>(div_round_up(words / 4 + 1, 256), 8192);
        hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, stream, d_table, words,
                           (int)(((uintptr_t)d_table & 15u) == 0));
        NLOD_LAUNCH_CHECK();
    }
    if (p.n_tri == 0) return;
    auto workspace = [&](size_t bytes) { return no_alloc ? fixed : tile_workspace(stream, bytes); };
    if (flags & NLOD_VOX_SOLID) launch_voxelize<true>(stream, p, geo, tiled, d_tris, d_table, workspace);
    else launch_voxelize<false>(stream, p, geo, tiled, d_tris, d_table, workspace);
}

void mesh_bbox(hipStream_t stream, const float* d_tris, uint64_t n_tri, float bmin[3], float bmax[3]) {
    NLOD_CHECK(n_tri > 0 && d_tris != nullptr, "mesh must have at least one triangle");
    uint32_t* acc = bbox_accumulator(stream);
    hipLaunchKernelGGL(init_bbox_kernel, dim3(1), dim3(64), 0, stream, acc);
    NLOD_LAUNCH_CHECK();
    const bool aligned16 = (reinterpret_cast<uintptr_t>(d_tris) & 15u) == 0;
    const uint64_t n_vals = n_tri * 9;
    const uint64_t per_block = aligned16 ? 12 * kBboxThreads : kBboxThreads;  // floats per block step
    const uint32_t blocks =
        (uint32_t)std::min<uint64_t>(div_round_up(n_vals, per_block), (uint64_t)cu_count() * 4);
    hipLaunchKernelGGL(bbox_kernel, dim3(blocks), dim3(kBboxThreads), 0, stream, d_tris, n_vals, aligned16, acc);
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
    return nlod::guarded([&] { nlod::voxelize(nlod::as_stream(stream), info, d_triangles, d_table, flags, false); });
}

int nlod_voxelize_noalloc(void* stream, const nlod_voxinfo* info, const float* d_triangles, uint32_t* d_table,
                          uint32_t flags) {
    return nlod::guarded([&] { nlod::voxelize(nlod::as_stream(stream), info, d_triangles, d_table, flags, true); });
}

int nlod_voxelize_mesh(void* stream, const float* d_triangles, uint64_t n_triangles, uint32_t res, uint32_t flags,
                       uint32_t* d_table, nlod_voxinfo* info_out) {
    return nlod::guarded([&] {
        float bmin[3], bmax[3];
        nlod::mesh_bbox(nlod::as_stream(stream), d_triangles, n_triangles, bmin, bmax);
        nlod_voxinfo info;
        nlod::voxinfo_from_bbox(bmin, bmax, res, flags, n_triangles, &info);
        nlod::voxelize(nlod::as_stream(stream), &info, d_triangles, d_table, flags, false);
        if (info_out) *info_out = info;
    });
}

}  // extern "C"
