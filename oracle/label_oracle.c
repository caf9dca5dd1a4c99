/*
 * ORACLE — test infrastructure only.  Nothing in the product links, loads or
 * calls this file; only tests/ use it, and only as the checker.
 *
 * CPU restatement of the training samples drawn from a mesh.  Visibility:
 * rays from the occupied voxels of a res^3 grid, labelled with the renderer's
 * visibility query (rptr/vulkan/raygen.rgen:690-730: nearest hit along the
 * ray, occlusion 1 when the hit point lies inside the voxel box).  Throughput:
 * the renderer's voxel-local throughput query (raygen.rgen:505-688) averaged
 * over spp paths at a level of the LoD pyramid (level l: (res >> l)^3 grid,
 * occupied where any level-0 voxel of its 2^l block is).  Specified to the bit
 * (the GPU kernels in neural_lod_amd/csrc/labels.hip follow the same steps):
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
 *   throughput  level and voxel by index draws, wo and wi on the sphere;
 *             per path an entry point by rejection of plane offsets in
 *             [-U, U]^2 around centre + 2 U wo (U the largest unit) until the
 *             ray along -wo meets the box (64 draws, else invalid); at up to
 *             max_depth hits inside the box: two-sided normal, next-event
 *             estimate T albedo / pi cos along wi when its shadow ray
 *             (t_min = 1e-4 (1 + |p|)) leaves the box, cosine bounce from a
 *             rejected unit-disk point in the Duff et al. basis, T *= albedo;
 *             target = mean over valid paths (first ray hits in the box)
 *   albedo    0.1 + 0.8 u per component, u from splitmix64 seeded
 *             albedo_seed ^ (triangle * 0x9E3779B97F4A7C15)
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

typedef struct {
    uint32_t res, n_occ;
    float unit[3];
    uint32_t* occ;
    uint64_t* start;
    uint32_t* lst;
} olevel;

static void free_levels(olevel* lv, uint32_t n) {
    for (uint32_t l = 0; l < n; ++l) {
        free(lv[l].occ);
        free(lv[l].start);
        free(lv[l].lst);
    }
}

/* Occupied voxels and triangle lists of every level; 0 when level 0 is empty. */
static int build_levels(const float* tris, uint64_t n_tri, uint32_t res, uint32_t n_lods, nlo_voxinfo* info,
                        olevel* lv) {
    float bmin[3], bmax[3];
    nlo_mesh_bbox(tris, n_tri, bmin, bmax);
    nlo_voxinfo_from_bbox(bmin, bmax, res, 0.0f, n_tri, info);
    const uint64_t n_vox0 = (uint64_t)res * res * res;
    uint32_t* bits = (uint32_t*)calloc((n_vox0 + 31) / 32, sizeof(uint32_t));
    nlo_voxelize_surface(info, tris, bits, 0);
    memset(lv, 0, n_lods * sizeof(olevel));
    int ok = 1;
    for (uint32_t l = 0; l < n_lods && ok; ++l) {
        const uint32_t R = res >> l;
        olevel* L = &lv[l];
        L->res = R;
        for (int k = 0; k < 3; ++k) L->unit[k] = (info->bbox_max[k] - info->bbox_min[k]) / (float)R;
        const uint64_t n_vox = (uint64_t)R * R * R;
        uint8_t* occ = (uint8_t*)calloc(n_vox, 1);
        for (uint32_t z = 0; z < res; ++z)
            for (uint32_t y = 0; y < res; ++y)
                for (uint32_t x = 0; x < res; ++x) {
                    uint64_t v = (uint64_t)x + (uint64_t)y * res + (uint64_t)z * res * res;
                    if (bits[v >> 5] & (1u << (31 - (v & 31))))
                        occ[(uint64_t)(x >> l) + (uint64_t)(y >> l) * R + (uint64_t)(z >> l) * R * R] = 1;
                }
        int64_t* slot_of = (int64_t*)malloc(n_vox * sizeof(int64_t));
        L->occ = (uint32_t*)malloc(n_vox * sizeof(uint32_t));
        for (uint64_t v = 0; v < n_vox; ++v) {
            slot_of[v] = -1;
            if (occ[v]) {
                slot_of[v] = L->n_occ;
                L->occ[L->n_occ++] = (uint32_t)v;
            }
        }
        free(occ);
        if (L->n_occ == 0) {
            ok = 0;
            free(slot_of);
            break;
        }
        L->start = (uint64_t*)calloc((size_t)L->n_occ + 1, sizeof(uint64_t));
        for (int pass = 0; pass < 2; ++pass) {
            uint64_t* fill = NULL;
            if (pass == 1) {
                for (uint32_t k = 1; k <= L->n_occ; ++k) L->start[k] += L->start[k - 1];
                L->lst = (uint32_t*)malloc((L->start[L->n_occ] + 1) * sizeof(uint32_t));
                fill = (uint64_t*)malloc((size_t)L->n_occ * sizeof(uint64_t));
                memcpy(fill, L->start, (size_t)L->n_occ * sizeof(uint64_t));
            }
            for (uint64_t t = 0; t < n_tri; ++t) {
                const float* tri = tris + t * 9;
                if (!tri_ok(tri)) continue;
                int lo[3], hi[3];
                for (int k = 0; k < 3; ++k) axis_range(tri, k, info->bbox_min[k], L->unit[k], R, &lo[k], &hi[k]);
                for (int z = lo[2]; z <= hi[2]; ++z)
                    for (int y = lo[1]; y <= hi[1]; ++y)
                        for (int x = lo[0]; x <= hi[0]; ++x) {
                            int64_t sl = slot_of[(uint64_t)x + (uint64_t)y * R + (uint64_t)z * R * R];
                            if (sl < 0) continue;
                            if (pass == 0) L->start[sl + 1]++;
                            else L->lst[fill[sl]++] = (uint32_t)t;
                        }
            }
            free(fill);
        }
        free(slot_of);
    }
    free(bits);
    return ok;
}

static uint32_t index_draw(uint64_t* s, uint32_t m) { return (uint32_t)(((splitmix64(s) >> 32) * (uint64_t)m) >> 32); }

static lv3 sphere_dir(uint64_t* st) {
    float a = 0.0f, b = 0.0f, c = 1.0f, s = 1.0f;
    for (int k = 0; k < 64; ++k) {
        float a1 = uniform01(st) * 2.0f - 1.0f;
        float b1 = uniform01(st) * 2.0f - 1.0f;
        float c1 = uniform01(st) * 2.0f - 1.0f;
        float s1 = (a1 * a1 + b1 * b1) + c1 * c1;
        if (s1 > 0.0f && s1 <= 1.0f) {
            a = a1; b = b1; c = c1; s = s1;
            break;
        }
    }
    float r = sqrtf(s);
    lv3 d = {a / r, b / r, c / r};
    return d;
}

/* nearest hit with t >= t_min: the voxel's list, or every finite triangle */
static float nearest_hit(const olevel* L, uint32_t slot, const float* tris, uint64_t n_tri, int brute_force, lv3 o,
                         lv3 d, float t_min, uint32_t* tri_out) {
    float best = INFINITY;
    if (brute_force) {
        for (uint64_t t = 0; t < n_tri; ++t) {
            if (!tri_ok(tris + t * 9)) continue;
            float th = ray_triangle(o, d, tris + t * 9);
            if (th >= t_min && th < best) { best = th; *tri_out = (uint32_t)t; }
        }
    } else {
        for (uint64_t k = L->start[slot]; k < L->start[slot + 1]; ++k) {
            float th = ray_triangle(o, d, tris + (uint64_t)L->lst[k] * 9);
            if (th >= t_min && th < best) { best = th; *tri_out = L->lst[k]; }
        }
    }
    return best;
}

static lv3 along(lv3 o, float t, lv3 d) { lv3 r = {o.x + t * d.x, o.y + t * d.y, o.z + t * d.z}; return r; }

typedef struct { lv3 lo, hi, c, h; } obox;

static obox voxel_box(const nlo_voxinfo* info, const olevel* L, uint32_t vx, uint32_t vy, uint32_t vz) {
    obox b;
    const float* bm = info->bbox_min;
    b.lo.x = bm[0] + (float)vx * L->unit[0]; b.lo.y = bm[1] + (float)vy * L->unit[1]; b.lo.z = bm[2] + (float)vz * L->unit[2];
    b.hi.x = bm[0] + (float)(vx + 1) * L->unit[0]; b.hi.y = bm[1] + (float)(vy + 1) * L->unit[1];
    b.hi.z = bm[2] + (float)(vz + 1) * L->unit[2];
    b.c.x = bm[0] + ((float)vx + 0.5f) * L->unit[0]; b.c.y = bm[1] + ((float)vy + 0.5f) * L->unit[1];
    b.c.z = bm[2] + ((float)vz + 0.5f) * L->unit[2];
    b.h.x = 0.5f * L->unit[0]; b.h.y = 0.5f * L->unit[1]; b.h.z = 0.5f * L->unit[2];
    return b;
}

static int in_box(const obox* b, lv3 p) {
    return fabsf(p.x - b->c.x) <= b->h.x && fabsf(p.y - b->c.y) <= b->h.y && fabsf(p.z - b->c.z) <= b->h.z;
}

/* Returns the number of occupied level-0 voxels (0: none, the outputs are untouched). */
int64_t nlo_visibility_samples(const float* tris, uint64_t n_tri, uint32_t res, uint64_t seed, uint64_t first,
                               uint32_t n, int brute_force, float* inputs, float* labels) {
    nlo_voxinfo info;
    olevel lv[1];
    if (!build_levels(tris, n_tri, res, 1, &info, lv)) {
        free_levels(lv, 1);
        return 0;
    }
    const olevel* L = &lv[0];
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t st = seed ^ ((first + i) * 0x9E3779B97F4A7C15ull);
        const uint32_t slot = index_draw(&st, L->n_occ);
        const uint32_t lin = L->occ[slot];
        const uint32_t vx = lin % res, vy = (lin / res) % res, vz = lin / (res * res);
        const float ux = uniform01(&st), uy = uniform01(&st), uz = uniform01(&st);
        lv3 o = {info.bbox_min[0] + ((float)vx + ux) * L->unit[0], info.bbox_min[1] + ((float)vy + uy) * L->unit[1],
                 info.bbox_min[2] + ((float)vz + uz) * L->unit[2]};
        lv3 d = sphere_dir(&st);
        uint32_t tri = 0;
        float best = nearest_hit(L, slot, tris, n_tri, brute_force, o, d, 0.0f, &tri);
        float label = 0.0f;
        if (best < INFINITY) {
            obox b = voxel_box(&info, L, vx, vy, vz);
            label = in_box(&b, along(o, best, d)) ? 1.0f : 0.0f;
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
    const int64_t n_occ = L->n_occ;
    free_levels(lv, 1);
    return n_occ;
}

static lv3 albedo_of(uint64_t albedo_seed, uint32_t tri) {
    uint64_t s = albedo_seed ^ ((uint64_t)tri * 0x9E3779B97F4A7C15ull);
    float a = uniform01(&s), b = uniform01(&s), c = uniform01(&s);
    lv3 r = {0.1f + 0.8f * a, 0.1f + 0.8f * b, 0.1f + 0.8f * c};
    return r;
}

/* Duff et al. 2017, "Building an Orthonormal Basis, Revisited" */
static void onb(lv3 n, lv3* b1, lv3* b2) {
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    b1->x = 1.0f + sign * n.x * n.x * a; b1->y = sign * b; b1->z = -sign * n.x;
    b2->x = b; b2->y = sign + n.y * n.y * a; b2->z = -n.y;
}

static float ray_eps(lv3 o) { return 1e-4f * (1.0f + sqrtf(ldot(o, o))); }

/* Throughput samples; n_occ_out[l] (nullable) = occupied voxels of level l.
 * Returns 0 when a level is empty (outputs untouched), else 1. */
int nlo_throughput_samples(const float* tris, uint64_t n_tri, uint32_t res, uint32_t n_lods, uint64_t seed,
                           uint64_t first, uint32_t n, uint32_t spp, uint32_t max_depth, uint64_t albedo_seed,
                           int brute_force, float* inputs, float* targets, float* valid, uint32_t* n_occ_out) {
    nlo_voxinfo info;
    olevel lv[8];
    if (n_lods < 1 || n_lods > 8 || !build_levels(tris, n_tri, res, n_lods, &info, lv)) {
        if (n_lods >= 1 && n_lods <= 8) free_levels(lv, n_lods);
        return 0;
    }
    if (n_occ_out)
        for (uint32_t l = 0; l < n_lods; ++l) n_occ_out[l] = lv[l].n_occ;
    const float kInvPi = 0.318309886183790671538f;
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t st = seed ^ ((first + i) * 0x9E3779B97F4A7C15ull);
        const uint32_t lod = index_draw(&st, n_lods);
        const olevel* L = &lv[lod];
        const uint32_t slot = index_draw(&st, L->n_occ);
        const uint32_t lin = L->occ[slot], R = L->res;
        const uint32_t vx = lin % R, vy = (lin / R) % R, vz = lin / (R * R);
        const lv3 wo = sphere_dir(&st);
        const lv3 wi = sphere_dir(&st);
        const obox box = voxel_box(&info, L, vx, vy, vz);
        const float U = fmaxf(L->unit[0], fmaxf(L->unit[1], L->unit[2]));
        lv3 u_ax, v_ax;
        onb(wo, &u_ax, &v_ax);
        const lv3 d0 = {-wo.x, -wo.y, -wo.z};
        const float back = 2.0f * U;
        lv3 sum = {0.0f, 0.0f, 0.0f};
        uint32_t n_valid = 0;
        for (uint32_t j = 0; j < spp; ++j) {
            int found = 0;
            lv3 entry = box.c;
            for (int k = 0; k < 64 && !found; ++k) {
                const float ox = (uniform01(&st) * 2.0f - 1.0f) * U;
                const float oy = (uniform01(&st) * 2.0f - 1.0f) * U;
                lv3 o = {((box.c.x + wo.x * back) + u_ax.x * ox) + v_ax.x * oy,
                         ((box.c.y + wo.y * back) + u_ax.y * ox) + v_ax.y * oy,
                         ((box.c.z + wo.z * back) + u_ax.z * ox) + v_ax.z * oy};
                float tn = -INFINITY, tf = INFINITY;
                int miss = 0;
                const float oc[3] = {o.x, o.y, o.z}, dc[3] = {d0.x, d0.y, d0.z};
                const float lo[3] = {box.lo.x, box.lo.y, box.lo.z}, hi[3] = {box.hi.x, box.hi.y, box.hi.z};
                for (int a = 0; a < 3; ++a) {
                    if (dc[a] == 0.0f) {
                        if (oc[a] < lo[a] || oc[a] > hi[a]) miss = 1;
                        continue;
                    }
                    float t0 = (lo[a] - oc[a]) / dc[a], t1 = (hi[a] - oc[a]) / dc[a];
                    if (t1 < t0) { float tmp = t0; t0 = t1; t1 = tmp; }
                    if (t0 > tn) tn = t0;
                    if (t1 < tf) tf = t1;
                }
                if (!miss && tn <= tf && tf >= 0.0f) {
                    entry = along(o, tn, d0);
                    found = 1;
                }
            }
            if (!found) continue;
            lv3 o = entry, d = d0, T = {1.0f, 1.0f, 1.0f}, acc = {0.0f, 0.0f, 0.0f};
            float t_min = 0.0f;
            int path_valid = 0;
            for (uint32_t bounce = 0; bounce < max_depth; ++bounce) {
                uint32_t tri = 0;
                const float t = nearest_hit(L, slot, tris, n_tri, brute_force, o, d, t_min, &tri);
                if (!(t < INFINITY)) break;
                const lv3 p = along(o, t, d);
                if (!in_box(&box, p)) break;
                if (bounce == 0) path_valid = 1;
                const float* tv = tris + (uint64_t)tri * 9;
                lv3 e1 = {tv[3] - tv[0], tv[4] - tv[1], tv[5] - tv[2]};
                lv3 e2 = {tv[6] - tv[0], tv[7] - tv[1], tv[8] - tv[2]};
                lv3 nrm = lcross(e1, e2);
                const float inv_len = 1.0f / sqrtf(ldot(nrm, nrm));
                nrm.x = nrm.x * inv_len; nrm.y = nrm.y * inv_len; nrm.z = nrm.z * inv_len;
                const lv3 wout = {-d.x, -d.y, -d.z};
                if (ldot(wout, nrm) < 0.0f) { nrm.x = -nrm.x; nrm.y = -nrm.y; nrm.z = -nrm.z; }
                const lv3 alb = albedo_of(albedo_seed, tri);
                const float eps = ray_eps(p);
                const float cos_i = ldot(wi, nrm);
                if (cos_i > 0.0f) {
                    uint32_t t2i = 0;
                    const float t2 = nearest_hit(L, slot, tris, n_tri, brute_force, p, wi, eps, &t2i);
                    if (!(t2 < INFINITY) || !in_box(&box, along(p, t2, wi))) {
                        acc.x = acc.x + (T.x * (alb.x * kInvPi)) * cos_i;
                        acc.y = acc.y + (T.y * (alb.y * kInvPi)) * cos_i;
                        acc.z = acc.z + (T.z * (alb.z * kInvPi)) * cos_i;
                    }
                }
                float x = 0.0f, y = 0.0f;
                for (int k = 0; k < 64; ++k) {
                    const float x1 = uniform01(&st) * 2.0f - 1.0f, y1 = uniform01(&st) * 2.0f - 1.0f;
                    if (x1 * x1 + y1 * y1 <= 1.0f) { x = x1; y = y1; break; }
                }
                const float zz = sqrtf(fmaxf(0.0f, 1.0f - (x * x + y * y)));
                lv3 b1, b2;
                onb(nrm, &b1, &b2);
                d.x = (b1.x * x + b2.x * y) + nrm.x * zz;
                d.y = (b1.y * x + b2.y * y) + nrm.y * zz;
                d.z = (b1.z * x + b2.z * y) + nrm.z * zz;
                T.x = T.x * alb.x; T.y = T.y * alb.y; T.z = T.z * alb.z;
                o = p;
                t_min = eps;
            }
            if (path_valid) {
                sum.x = sum.x + acc.x; sum.y = sum.y + acc.y; sum.z = sum.z + acc.z;
                ++n_valid;
            }
        }
        float* in = inputs + (uint64_t)i * 10;
        in[0] = (box.c.x - info.bbox_min[0]) / (info.bbox_max[0] - info.bbox_min[0]);
        in[1] = (box.c.y - info.bbox_min[1]) / (info.bbox_max[1] - info.bbox_min[1]);
        in[2] = (box.c.z - info.bbox_min[2]) / (info.bbox_max[2] - info.bbox_min[2]);
        in[3] = (wo.x + 1.0f) * 0.5f; in[4] = (wo.y + 1.0f) * 0.5f; in[5] = (wo.z + 1.0f) * 0.5f;
        in[6] = (wi.x + 1.0f) * 0.5f; in[7] = (wi.y + 1.0f) * 0.5f; in[8] = (wi.z + 1.0f) * 0.5f;
        in[9] = (float)lod;
        const float nv = (float)n_valid;
        float* out = targets + (uint64_t)i * 3;
        out[0] = n_valid ? sum.x / nv : 0.0f;
        out[1] = n_valid ? sum.y / nv : 0.0f;
        out[2] = n_valid ? sum.z / nv : 0.0f;
        valid[i] = n_valid ? 1.0f : 0.0f;
    }
    free_levels(lv, n_lods);
    return 1;
}
