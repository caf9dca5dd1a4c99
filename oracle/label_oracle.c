/*
 * ORACLE — test infrastructure only.  Nothing in the product links, loads or
 * calls this file; only tests/ use it, and only as the checker.
 *
 * CPU restatement of the visibility training samples: rays from the occupied
 * voxels of a res^3 grid over a mesh, labelled with the renderer's visibility
 * query (rptr/vulkan/raygen.rgen:690-730: nearest hit along the ray, occlusion
 * 1 when the hit point lies inside the voxel box).  The sampling is the
 * visibility training configuration's (voxel uniform over the occupied ones,
 * origin uniform in it, direction uniform on the sphere), specified to the bit:
 *
 *   grid      nlo_voxinfo_from_bbox(mesh box, res, no pad), occupancy from
 *             nlo_voxelize_surface, occupied voxels in ascending linear index
 *   lists     per occupied voxel the finite triangles whose clamped voxel range
 *             (int)clamp((min/max vertex - bmin) / unit, 0, res-1) contains it
 *             (cpu_voxelizer.cpp:70-84); brute_force = every finite triangle
 *   rng       state = seed ^ (index * 0x9E3779B97F4A7C15); splitmix64;
 *             u = (x >> 40) * 2^-24
 *   voxel     slot = ((x >> 32) * n_occupied) >> 32
 *   origin    bmin + ((float)v + u) * unit per axis
 *   direction rejection in [-1,1]^3 until 0 < s <= 1 (64 draws at most, then
 *             (0,0,1)); (a,b,c) / sqrtf(s)
 *   hit       Moller-Trumbore, nearest t >= 0; label 1 when every
 *             |p - centre| <= unit / 2
 *   inputs    (o - bmin) / (bmax - bmin), (d + 1) / 2
 *
 * Built with -ffp-contract=off: every float operation is rounded on its own.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float bbox_min[3];
    float bbox_max[3];
    uint32_t gridsize[3];
    uint32_t pad0;
    uint64_t n_triangles;
    float unit[3];
    uint32_t pad1;
} nlo_voxinfo;

void nlo_mesh_bbox(const float* tris, uint64_t n_tri, float out_min[3], float out_max[3]);
void nlo_voxinfo_from_bbox(const float bmin[3], const float bmax[3], uint32_t G, float eps_pad, uint64_t n_tri,
                           nlo_voxinfo* info);
void nlo_voxelize_surface(const nlo_voxinfo* info, const float* tris, uint32_t* table, int morton);

typedef struct { float x, y, z; } lv3;

static uint64_t splitmix64(uint64_t* s) {
    *s += 0x9E3779B97F4A7C15ull;
    uint64_t z = *s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static float uniform01(uint64_t* s) { return (float)(splitmix64(s) >> 40) * 0x1p-24f; }

static lv3 lsub(lv3 a, lv3 b) { lv3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static float ldot(lv3 a, lv3 b) { float t0 = a.x * b.x, t1 = a.y * b.y, t2 = a.z * b.z; return (t0 + t1) + t2; }
static lv3 lcross(lv3 a, lv3 b) {
    lv3 r = {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
    return r;
}

/* Moller-Trumbore (Moller & Trumbore 1997): t of the hit, negative or NaN for none. */
static float ray_triangle(lv3 o, lv3 d, const float* tri) {
    lv3 p0 = {tri[0], tri[1], tri[2]}, p1 = {tri[3], tri[4], tri[5]}, p2 = {tri[6], tri[7], tri[8]};
    lv3 e1 = lsub(p1, p0), e2 = lsub(p2, p0);
    lv3 pv = lcross(d, e2);
    float det = ldot(e1, pv);
    if (det == 0.0f) return -1.0f;
    float inv = 1.0f / det;
    lv3 s = lsub(o, p0);
    float u = ldot(s, pv) * inv;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    lv3 q = lcross(s, e1);
    float w = ldot(d, q) * inv;
    if (w < 0.0f || u + w > 1.0f) return -1.0f;
    return ldot(e2, q) * inv;
}

static float gmin_f(float x, float y) { return (y < x) ? y : x; }
static float gmax_f(float x, float y) { return (x < y) ? y : x; }

static int tri_ok(const float* t) {
    for (int k = 0; k < 9; ++k)
        if (!isfinite(t[k])) return 0;
    return 1;
}

static void axis_range(const float* tri, int k, float bmin, float unit, uint32_t G, int* lo, int* hi) {
    float a = tri[k] - bmin, b = tri[3 + k] - bmin, c = tri[6 + k] - bmin;
    float gm = (float)(G - 1);
    *lo = (int)gmin_f(gmax_f(gmin_f(a, gmin_f(b, c)) / unit, 0.0f), gm);
    *hi = (int)gmin_f(gmax_f(gmax_f(a, gmax_f(b, c)) / unit, 0.0f), gm);
}

/* Returns the number of occupied voxels (0: none, the outputs are untouched). */
int64_t nlo_visibility_samples(const float* tris, uint64_t n_tri, uint32_t res, uint64_t seed, uint64_t first,
                               uint32_t n, int brute_force, float* inputs, float* labels) {
    float bmin[3], bmax[3];
    nlo_mesh_bbox(tris, n_tri, bmin, bmax);
    nlo_voxinfo info;
    nlo_voxinfo_from_bbox(bmin, bmax, res, 0.0f, n_tri, &info);
    const uint64_t n_vox = (uint64_t)res * res * res;
    uint32_t* bits = (uint32_t*)calloc((n_vox + 31) / 32, sizeof(uint32_t));
    nlo_voxelize_surface(&info, tris, bits, 0);
    uint32_t* occ = (uint32_t*)malloc(n_vox * sizeof(uint32_t));
    int64_t* slot_of = (int64_t*)malloc(n_vox * sizeof(int64_t));
    uint32_t n_occ = 0;
    for (uint64_t v = 0; v < n_vox; ++v) {
        slot_of[v] = -1;
        if (bits[v >> 5] & (1u << (31 - (v & 31)))) {
            slot_of[v] = n_occ;
            occ[n_occ++] = (uint32_t)v;
        }
    }
    free(bits);
    if (n_occ == 0) {
        free(occ);
        free(slot_of);
        return 0;
    }
    /* per-slot triangle lists (CSR) */
    uint64_t* start = (uint64_t*)calloc((size_t)n_occ + 1, sizeof(uint64_t));
    uint32_t* lst = NULL;
    for (int pass = 0; pass < 2; ++pass) {
        uint64_t* fill = NULL;
        if (pass == 1) {
            for (uint32_t k = 1; k <= n_occ; ++k) start[k] += start[k - 1];
            lst = (uint32_t*)malloc((start[n_occ] + 1) * sizeof(uint32_t));
            fill = (uint64_t*)malloc((size_t)n_occ * sizeof(uint64_t));
            memcpy(fill, start, (size_t)n_occ * sizeof(uint64_t));
        }
        for (uint64_t t = 0; t < n_tri && !brute_force; ++t) {
            const float* tri = tris + t * 9;
            if (!tri_ok(tri)) continue;
            int lo[3], hi[3];
            for (int k = 0; k < 3; ++k) axis_range(tri, k, info.bbox_min[k], info.unit[k], res, &lo[k], &hi[k]);
            for (int z = lo[2]; z <= hi[2]; ++z)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int x = lo[0]; x <= hi[0]; ++x) {
                        int64_t sl = slot_of[(uint64_t)x + (uint64_t)y * res + (uint64_t)z * res * res];
                        if (sl < 0) continue;
                        if (pass == 0) start[sl + 1]++;
                        else lst[fill[sl]++] = (uint32_t)t;
                    }
        }
        free(fill);
    }
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t st = seed ^ ((first + i) * 0x9E3779B97F4A7C15ull);
        const uint32_t slot = (uint32_t)(((splitmix64(&st) >> 32) * (uint64_t)n_occ) >> 32);
        const uint32_t lin = occ[slot];
        const uint32_t vx = lin % res, vy = (lin / res) % res, vz = lin / (res * res);
        const float ux = uniform01(&st), uy = uniform01(&st), uz = uniform01(&st);
        lv3 o = {info.bbox_min[0] + ((float)vx + ux) * info.unit[0], info.bbox_min[1] + ((float)vy + uy) * info.unit[1],
                 info.bbox_min[2] + ((float)vz + uz) * info.unit[2]};
        float a = 0.0f, b = 0.0f, c = 1.0f, s = 1.0f;
        for (int k = 0; k < 64; ++k) {
            float a1 = uniform01(&st) * 2.0f - 1.0f;
            float b1 = uniform01(&st) * 2.0f - 1.0f;
            float c1 = uniform01(&st) * 2.0f - 1.0f;
            float s1 = (a1 * a1 + b1 * b1) + c1 * c1;
            if (s1 > 0.0f && s1 <= 1.0f) {
                a = a1; b = b1; c = c1; s = s1;
                break;
            }
        }
        float r = sqrtf(s);
        lv3 d = {a / r, b / r, c / r};
        float best = INFINITY;
        if (brute_force) {
            for (uint64_t t = 0; t < n_tri; ++t) {
                if (!tri_ok(tris + t * 9)) continue;
                float th = ray_triangle(o, d, tris + t * 9);
                if (th >= 0.0f && th < best) best = th;
            }
        } else {
            for (uint64_t k = start[slot]; k < start[slot + 1]; ++k) {
                float th = ray_triangle(o, d, tris + (uint64_t)lst[k] * 9);
                if (th >= 0.0f && th < best) best = th;
            }
        }
        float label = 0.0f;
        if (best < INFINITY) {
            lv3 p = {o.x + best * d.x, o.y + best * d.y, o.z + best * d.z};
            lv3 cc = {info.bbox_min[0] + ((float)vx + 0.5f) * info.unit[0],
                      info.bbox_min[1] + ((float)vy + 0.5f) * info.unit[1],
                      info.bbox_min[2] + ((float)vz + 0.5f) * info.unit[2]};
            int inside = fabsf(p.x - cc.x) <= 0.5f * info.unit[0] && fabsf(p.y - cc.y) <= 0.5f * info.unit[1] &&
                         fabsf(p.z - cc.z) <= 0.5f * info.unit[2];
            label = inside ? 1.0f : 0.0f;
        }
        float* in = inputs + (uint64_t)i * 6;
        in[0] = (o.x - info.bbox_min[0]) / (info.bbox_max[0] - info.bbox_min[0]);
        in[1] = (o.y - info.bbox_min[1]) / (info.bbox_max[1] - info.bbox_min[1]);
        in[2] = (o.z - info.bbox_min[2]) / (info.bbox_max[2] - info.bbox_min[2]);
        in[3] = (d.x + 1.0f) * 0.5f;
        in[4] = (d.y + 1.0f) * 0.5f;
        in[5] = (d.z + 1.0f) * 0.5f;
        labels[i] = label;
    }
    free(start);
    free(lst);
    free(occ);
    free(slot_of);
    return n_occ;
}
