/*
 * ORACLE — test infrastructure only.  Nothing in the product links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker.
 *
 * CPU restatement of the reference voxelizer (cuda_voxelizer v0.5.2 as vendored
 * in neural_lod), written in plain C with IEEE single precision per operation:
 * build with -ffp-contract=off so that no multiply-add is fused.
 *
 *   nlo_voxinfo_from_bbox  <- util.h:63-68 (voxinfo ctor) + util.h:87-112
 *                            (createMeshBBCube, optional epsilon pad) and
 *                            main.cpp:209-216 (pad 1e-4 when -pad is given)
 *   nlo_mesh_bbox          <- trimesh2/libsrc/TriMesh_bounding.cc:18-31
 *                            (per-vertex min/max)
 *   nlo_voxelize_surface   <- cpu_voxelizer.cpp:39-193 (cpu_voxelize_mesh)
 *   nlo_voxelize_solid     <- cpu_voxelizer.cpp:195-328 (cpu_voxelize_mesh_solid)
 *
 * The vector arithmetic of the reference goes through GLM 0.9.9.8, which is
 * not vendored in the reference.  Its published semantics are spelled out
 * here operation by operation:
 *   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z        (vec2: a.x*b.x + a.y*b.y)
 *   cross(a,b)    = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
 *   normalize(v)  = v * (1 / sqrt(dot(v,v)))
 *   min/max/clamp = componentwise, clamp(x,lo,hi) = min(max(x,lo),hi)
 *   vec3 -> ivec3 = truncation toward zero
 * Parity unpinned for this operation order: the reference's CPU path cannot be
 * compiled here (GLM and trimesh2 are absent) and its repository holds no
 * voxelized output.  The geometry is pinned (separating-axis brute force on
 * small meshes, closed-form box shells, analytic solids); the rounding of
 * voxels whose test sits within float rounding of a boundary is not.
 *
 * Bit layout (util.h:32-45, cpu_voxelizer.cpp:9-17): voxel (x,y,z) has linear
 * index i = x + y*G + z*G*G and lives in bit 31-(i%32) of 32-bit word i/32.
 *
 * Places where the reference's behaviour is undefined are given a defined
 * result here and in the HIP kernels alike (both keep every access inside
 * the table):
 *   - a triangle with a non-finite vertex coordinate is skipped (the
 *     reference converts NaN to int when it clamps the triangle's box);
 *   - solid mode: the x-run end xmax is clamped to G-1 and a negative xmax
 *     sets nothing (the reference CPU path stores it in an unsigned int, so a
 *     negative value loops past the table; its GPU path writes past the end
 *     of a row when xmax >= G); the y/z sample ranges are clamped to [0,G-1]
 *     (a no-op whenever the mesh lies in the bounding box);
 *   - solid mode: a triangle whose normal is not finite (zero area) is
 *     skipped (the reference would convert NaN to int).
 * Morton output order (cpu_voxelizer.cpp:22-36) is supported as well.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* cpu_voxelizer.cpp:4 defines float_error as the double literal 0.000001, so
 * every comparison against it is made in double precision. */
#define FLOAT_ERROR 0.000001

typedef struct { float x, y, z; } f3;
typedef struct { float x, y; } f2;

typedef struct {
    float bbox_min[3];
    float bbox_max[3];
    uint32_t gridsize[3];
    uint32_t pad0;
    uint64_t n_triangles;
    float unit[3];
    uint32_t pad1;
} nlo_voxinfo;

static inline f3 v3_sub(f3 a, f3 b) { f3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline float v3_dot(f3 a, f3 b) { float t0 = a.x * b.x, t1 = a.y * b.y, t2 = a.z * b.z; return (t0 + t1) + t2; }
static inline float v2_dot(f2 a, f2 b) { float t0 = a.x * b.x, t1 = a.y * b.y; return t0 + t1; }
static inline f3 v3_cross(f3 a, f3 b) {
    f3 r;
    r.x = a.y * b.z - b.y * a.z;
    r.y = a.z * b.x - b.z * a.x;
    r.z = a.x * b.y - b.x * a.y;
    return r;
}
static inline f3 v3_normalize(f3 v) {
    float inv = 1.0f / sqrtf(v3_dot(v, v));
    f3 r = {v.x * inv, v.y * inv, v.z * inv};
    return r;
}
/* GLM min(x,y) = (y < x) ? y : x ; max(x,y) = (x < y) ? y : x */
static inline float gmin(float x, float y) { return (y < x) ? y : x; }
static inline float gmax(float x, float y) { return (x < y) ? y : x; }
static inline float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

/* Morton encoding: bit-interleave x (bit 0), y (bit 1), z (bit 2), as the
 * 256-entry lookup tables of morton_LUTs.h produce for coordinates < 2^21. */
static uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) {
        r |= (uint64_t)((x >> b) & 1u) << (3 * b);
        r |= (uint64_t)((y >> b) & 1u) << (3 * b + 1);
        r |= (uint64_t)((z >> b) & 1u) << (3 * b + 2);
    }
    return r;
}

static inline void set_bit_or(uint32_t* table, uint64_t index) {
    uint64_t w = index / 32u;
    uint32_t mask = 1u << (31u - (uint32_t)(index % 32u));
#pragma omp critical
    { table[w] |= mask; }
}

static inline void set_bit_xor(uint32_t* table, uint64_t index) {
    uint64_t w = index / 32u;
    uint32_t mask = 1u << (31u - (uint32_t)(index % 32u));
#pragma omp critical
    { table[w] ^= mask; }
}

/* Per-vertex bounding box of a triangle soup [n][3 vertices][xyz]. */
void nlo_mesh_bbox(const float* tris, uint64_t n_tri, float out_min[3], float out_max[3]) {
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t t = 0; t < n_tri; ++t) {
        for (int v = 0; v < 3; ++v) {
            for (int d = 0; d < 3; ++d) {
                float c = tris[t * 9 + v * 3 + d];
                if (c < mn[d]) mn[d] = c;
                if (c > mx[d]) mx[d] = c;
            }
        }
    }
    memcpy(out_min, mn, sizeof(mn));
    memcpy(out_max, mx, sizeof(mx));
}

/* createMeshBBCube + voxinfo constructor for a cubic G^3 grid. */
void nlo_voxinfo_from_bbox(const float bmin[3], const float bmax[3], uint32_t G, float eps_pad,
                           uint64_t n_tri, nlo_voxinfo* info) {
    float amin[3] = {bmin[0], bmin[1], bmin[2]};
    float amax[3] = {bmax[0], bmax[1], bmax[2]};
    float len[3] = {bmax[0] - bmin[0], bmax[1] - bmin[1], bmax[2] - bmin[2]};
    float maxlen = gmax(len[0], gmax(len[1], len[2]));
    for (int i = 0; i < 3; ++i) {
        if (maxlen == len[i]) continue;
        float delta = maxlen - len[i];
        amin[i] = bmin[i] - (delta / 2.0f);
        amax[i] = bmax[i] + (delta / 2.0f);
    }
    if (eps_pad != 0.0f) {
        float e[3];
        for (int i = 0; i < 3; ++i) e[i] = (amax[i] - amin[i]) * eps_pad;
        for (int i = 0; i < 3; ++i) { amin[i] = amin[i] - e[i]; amax[i] = amax[i] + e[i]; }
    }
    memset(info, 0, sizeof(*info));
    for (int i = 0; i < 3; ++i) {
        info->bbox_min[i] = amin[i];
        info->bbox_max[i] = amax[i];
        info->gridsize[i] = G;
        info->unit[i] = (amax[i] - amin[i]) / (float)G;
    }
    info->n_triangles = n_tri;
}

static inline int tri_finite(const float* tris, uint64_t t) {
    for (int k = 0; k < 9; ++k)
        if (!isfinite(tris[t * 9 + k])) return 0;
    return 1;
}

static inline f3 load_vertex(const float* tris, uint64_t t, int v, const nlo_voxinfo* info) {
    const float* p = tris + t * 9 + v * 3;
    /* The translate prepass (cpu_voxelizer.cpp:48-53) subtracts bbox.min in float. */
    f3 r = {p[0] - info->bbox_min[0], p[1] - info->bbox_min[1], p[2] - info->bbox_min[2]};
    return r;
}

static inline uint64_t linear_index(const nlo_voxinfo* info, int x, int y, int z, int morton) {
    if (morton) return morton3((uint32_t)x, (uint32_t)y, (uint32_t)z);
    /* cpu_voxelizer.cpp:178 uses gridsize.y and gridsize.z; the grid is cubic. */
    return (uint64_t)x + (uint64_t)y * (uint64_t)info->gridsize[1] +
           (uint64_t)z * (uint64_t)info->gridsize[1] * (uint64_t)info->gridsize[2];
}

/* cpu_voxelize_mesh (cpu_voxelizer.cpp:39-193): conservative surface voxelization. */
void nlo_voxelize_surface(const nlo_voxinfo* info_p, const float* tris, uint32_t* table, int morton) {
    const nlo_voxinfo info = *info_p;
    const int64_t n = (int64_t)info.n_triangles;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        if (!tri_finite(tris, (uint64_t)i)) continue;
        const f3 unit = {info.unit[0], info.unit[1], info.unit[2]};
        const f3 delta_p = unit;
        f3 c = {0.0f, 0.0f, 0.0f};
        const f3 grid_max = {(float)(info.gridsize[0] - 1), (float)(info.gridsize[1] - 1), (float)(info.gridsize[2] - 1)};

        f3 v0 = load_vertex(tris, (uint64_t)i, 0, &info);
        f3 v1 = load_vertex(tris, (uint64_t)i, 1, &info);
        f3 v2 = load_vertex(tris, (uint64_t)i, 2, &info);
        f3 e0 = v3_sub(v1, v0);
        f3 e1 = v3_sub(v2, v1);
        f3 e2 = v3_sub(v0, v2);
        f3 nrm = v3_normalize(v3_cross(e0, e1));

        f3 wmin = {gmin(v0.x, gmin(v1.x, v2.x)), gmin(v0.y, gmin(v1.y, v2.y)), gmin(v0.z, gmin(v1.z, v2.z))};
        f3 wmax = {gmax(v0.x, gmax(v1.x, v2.x)), gmax(v0.y, gmax(v1.y, v2.y)), gmax(v0.z, gmax(v1.z, v2.z))};
        int gmin_x = (int)gclamp(wmin.x / unit.x, 0.0f, grid_max.x);
        int gmin_y = (int)gclamp(wmin.y / unit.y, 0.0f, grid_max.y);
        int gmin_z = (int)gclamp(wmin.z / unit.z, 0.0f, grid_max.z);
        int gmax_x = (int)gclamp(wmax.x / unit.x, 0.0f, grid_max.x);
        int gmax_y = (int)gclamp(wmax.y / unit.y, 0.0f, grid_max.y);
        int gmax_z = (int)gclamp(wmax.z / unit.z, 0.0f, grid_max.z);

        if (nrm.x > 0.0f) c.x = unit.x;
        if (nrm.y > 0.0f) c.y = unit.y;
        if (nrm.z > 0.0f) c.z = unit.z;
        float d1 = v3_dot(nrm, v3_sub(c, v0));
        float d2 = v3_dot(nrm, v3_sub(v3_sub(delta_p, c), v0));

        /* XY projection */
        f2 nxy0 = {-1.0f * e0.y, e0.x}, nxy1 = {-1.0f * e1.y, e1.x}, nxy2 = {-1.0f * e2.y, e2.x};
        if (nrm.z < 0.0f) {
            nxy0.x = -nxy0.x; nxy0.y = -nxy0.y;
            nxy1.x = -nxy1.x; nxy1.y = -nxy1.y;
            nxy2.x = -nxy2.x; nxy2.y = -nxy2.y;
        }
        float dxy0 = ((-1.0f * v2_dot(nxy0, (f2){v0.x, v0.y})) + gmax(0.0f, unit.x * nxy0.x)) + gmax(0.0f, unit.y * nxy0.y);
        float dxy1 = ((-1.0f * v2_dot(nxy1, (f2){v1.x, v1.y})) + gmax(0.0f, unit.x * nxy1.x)) + gmax(0.0f, unit.y * nxy1.y);
        float dxy2 = ((-1.0f * v2_dot(nxy2, (f2){v2.x, v2.y})) + gmax(0.0f, unit.x * nxy2.x)) + gmax(0.0f, unit.y * nxy2.y);
        /* YZ projection */
        f2 nyz0 = {-1.0f * e0.z, e0.y}, nyz1 = {-1.0f * e1.z, e1.y}, nyz2 = {-1.0f * e2.z, e2.y};
        if (nrm.x < 0.0f) {
            nyz0.x = -nyz0.x; nyz0.y = -nyz0.y;
            nyz1.x = -nyz1.x; nyz1.y = -nyz1.y;
            nyz2.x = -nyz2.x; nyz2.y = -nyz2.y;
        }
        float dyz0 = ((-1.0f * v2_dot(nyz0, (f2){v0.y, v0.z})) + gmax(0.0f, unit.y * nyz0.x)) + gmax(0.0f, unit.z * nyz0.y);
        float dyz1 = ((-1.0f * v2_dot(nyz1, (f2){v1.y, v1.z})) + gmax(0.0f, unit.y * nyz1.x)) + gmax(0.0f, unit.z * nyz1.y);
        float dyz2 = ((-1.0f * v2_dot(nyz2, (f2){v2.y, v2.z})) + gmax(0.0f, unit.y * nyz2.x)) + gmax(0.0f, unit.z * nyz2.y);
        /* ZX projection.  The CPU path weights component 0 (a z quantity) by
         * unit.x and component 1 by unit.z (cpu_voxelizer.cpp:133-135); kept. */
        f2 nzx0 = {-1.0f * e0.x, e0.z}, nzx1 = {-1.0f * e1.x, e1.z}, nzx2 = {-1.0f * e2.x, e2.z};
        if (nrm.y < 0.0f) {
            nzx0.x = -nzx0.x; nzx0.y = -nzx0.y;
            nzx1.x = -nzx1.x; nzx1.y = -nzx1.y;
            nzx2.x = -nzx2.x; nzx2.y = -nzx2.y;
        }
        float dzx0 = ((-1.0f * v2_dot(nzx0, (f2){v0.z, v0.x})) + gmax(0.0f, unit.x * nzx0.x)) + gmax(0.0f, unit.z * nzx0.y);
        float dzx1 = ((-1.0f * v2_dot(nzx1, (f2){v1.z, v1.x})) + gmax(0.0f, unit.x * nzx1.x)) + gmax(0.0f, unit.z * nzx1.y);
        float dzx2 = ((-1.0f * v2_dot(nzx2, (f2){v2.z, v2.x})) + gmax(0.0f, unit.x * nzx2.x)) + gmax(0.0f, unit.z * nzx2.y);

        for (int z = gmin_z; z <= gmax_z; ++z) {
            for (int y = gmin_y; y <= gmax_y; ++y) {
                for (int x = gmin_x; x <= gmax_x; ++x) {
                    f3 p = {(float)x * unit.x, (float)y * unit.y, (float)z * unit.z};
                    float ndp = v3_dot(nrm, p);
                    if (((ndp + d1) * (ndp + d2)) > 0.0f) continue;
                    f2 pxy = {p.x, p.y};
                    if ((v2_dot(nxy0, pxy) + dxy0) < 0.0f) continue;
                    if ((v2_dot(nxy1, pxy) + dxy1) < 0.0f) continue;
                    if ((v2_dot(nxy2, pxy) + dxy2) < 0.0f) continue;
                    f2 pyz = {p.y, p.z};
                    if ((v2_dot(nyz0, pyz) + dyz0) < 0.0f) continue;
                    if ((v2_dot(nyz1, pyz) + dyz1) < 0.0f) continue;
                    if ((v2_dot(nyz2, pyz) + dyz2) < 0.0f) continue;
                    f2 pzx = {p.z, p.x};
                    if ((v2_dot(nzx0, pzx) + dzx0) < 0.0f) continue;
                    if ((v2_dot(nzx1, pzx) + dzx1) < 0.0f) continue;
                    if ((v2_dot(nzx2, pzx) + dzx2) < 0.0f) continue;
                    set_bit_or(table, linear_index(&info, x, y, z, morton));
                }
            }
        }
    }
}

/* check_point_triangle (cpu_voxelizer.cpp:228-249) */
static int check_point_triangle(f2 a, f2 b, f2 c, f2 p) {
    f2 pa = {p.x - a.x, p.y - a.y};
    f2 pb = {p.x - b.x, p.y - b.y};
    f2 pc = {p.x - c.x, p.y - c.y};
    float t1 = pa.x * pb.y - pa.y * pb.x;
    if ((double)fabsf(t1) < FLOAT_ERROR && pa.x * pb.x <= 0 && pa.y * pb.y <= 0) return 1;
    float t2 = pb.x * pc.y - pb.y * pc.x;
    if ((double)fabsf(t2) < FLOAT_ERROR && pb.x * pc.x <= 0 && pb.y * pc.y <= 0) return 2;
    float t3 = pc.x * pa.y - pc.y * pa.x;
    if ((double)fabsf(t3) < FLOAT_ERROR && pc.x * pa.x <= 0 && pc.y * pa.y <= 0) return 3;
    if (t1 * t2 > 0 && t1 * t3 > 0) return 0;
    return -1;
}

/* TopLeftEdge (cpu_voxelizer.cpp:206-208) */
static inline int top_left_edge(f2 a, f2 b) { return (b.y < a.y) || (b.y == a.y && a.x > b.x); }

/* checkCCW (cpu_voxelizer.cpp:211-219) */
static inline int check_ccw(f2 a, f2 b, f2 c) {
    f2 e0 = {b.x - a.x, b.y - a.y};
    f2 e1 = {c.x - a.x, c.y - a.y};
    float r = e0.x * e1.y - e1.x * e0.y;
    return r > 0;
}

/* cpu_voxelize_mesh_solid (cpu_voxelizer.cpp:252-328): YZ-projection parity fill. */
void nlo_voxelize_solid(const nlo_voxinfo* info_p, const float* tris, uint32_t* table, int morton) {
    const nlo_voxinfo info = *info_p;
    const int64_t n = (int64_t)info.n_triangles;
    const int G1 = (int)info.gridsize[1] - 1, G2 = (int)info.gridsize[2] - 1, G0 = (int)info.gridsize[0] - 1;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        if (!tri_finite(tris, (uint64_t)i)) continue;
        f3 v0 = load_vertex(tris, (uint64_t)i, 0, &info);
        f3 v1 = load_vertex(tris, (uint64_t)i, 1, &info);
        f3 v2 = load_vertex(tris, (uint64_t)i, 2, &info);
        f3 e0 = v3_sub(v1, v0);
        f3 e1 = v3_sub(v2, v1);
        f3 nrm = v3_normalize(v3_cross(e0, e1));
        if ((double)fabsf(nrm.x) < FLOAT_ERROR) continue;
        if (!(isfinite(nrm.x) && isfinite(nrm.y) && isfinite(nrm.z))) continue;

        f2 a = {v0.y, v0.z}, b = {v1.y, v1.z}, c = {v2.y, v2.z};
        if (!check_ccw(a, b, c)) { f2 t = b; b = c; c = t; }

        float bmax_x = gmax(a.x, gmax(b.x, c.x)), bmax_y = gmax(a.y, gmax(b.y, c.y));
        float bmin_x = gmin(a.x, gmin(b.x, c.x)), bmin_y = gmin(a.y, gmin(b.y, c.y));
        /* float quotient, then double arithmetic with the 0.5 literal, then back to float */
        float gmax_y = (float)floor((double)(bmax_x / info.unit[1]) - 0.5);
        float gmax_z = (float)floor((double)(bmax_y / info.unit[2]) - 0.5);
        float gmin_y = (float)ceil((double)(bmin_x / info.unit[1]) - 0.5);
        float gmin_z = (float)ceil((double)(bmin_y / info.unit[2]) - 0.5);
        /* float -> int truncation of the (clamped) sample range */
        int y0 = (int)fmaxf(gmin_y, -1.0f), y1 = (int)fminf(gmax_y, (float)info.gridsize[1]);
        int z0 = (int)fmaxf(gmin_z, -1.0f), z1 = (int)fminf(gmax_z, (float)info.gridsize[2]);
        if (y0 < 0) y0 = 0;
        if (z0 < 0) z0 = 0;
        if (y1 > G1) y1 = G1;
        if (z1 > G2) z1 = G2;

        for (int y = y0; y <= y1; ++y) {
            for (int z = z0; z <= z1; ++z) {
                f2 pt = {(float)(((double)y + 0.5) * (double)info.unit[1]),
                         (float)(((double)z + 0.5) * (double)info.unit[2])};
                int k = check_point_triangle(a, b, c, pt);
                if ((k == 1 && top_left_edge(a, b)) || (k == 2 && top_left_edge(b, c)) ||
                    (k == 3 && top_left_edge(c, a)) || (k == 0)) {
                    /* get_x_coordinate (cpu_voxelizer.cpp:222-224), float throughout */
                    float xc = (-(nrm.y * (pt.x - v0.y) + nrm.z * (pt.y - v0.z)) / nrm.x + v0.x);
                    double xq = (double)(xc / info.unit[0]) - 0.5;
                    if (!(xq > -1.0)) continue; /* int(xq) < 0: nothing to flip */
                    int xmax = (xq >= (double)G0) ? G0 : (int)xq;
                    for (int x = 0; x <= xmax; ++x) set_bit_xor(table, linear_index(&info, x, y, z, morton));
                }
            }
        }
    }
}

int nlo_omp_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void nlo_omp_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
