/*
 * ORACLE — test infrastructure only.  Nothing in the product links, loads or
 * calls this file; only tests/ use it, and only as the checker.
 *
 * CPU restatement of the LoD occupancy pyramid and the first-active-voxel
 * march (the traversal of first_active_voxel_intersection,
 * rptr/cuda/neural-lod/neural_lod_learning_common.cuh:784-832, over bit
 * tables; the GPU kernels in neural_lod_amd/csrc/lod_grid.hip follow the same
 * steps).  Bit layout: voxel i = x + y R + z R^2 in bit 31 - i % 32 of word
 * i / 32.  Built with -ffp-contract=off.
 *
 *   downsample  parent voxel occupied when any of its 2x2x2 children is
 *   march       p0 = (o - bmin) / unit, q = d / unit; slab clip to [0, R]^3
 *               from max(t_min, entry); start voxel clamp(floor(p0 + t q));
 *               per step: occupied and not prev -> voxel box slabs give
 *               t_near <= t_far -> hit; else advance the axis with the nearest
 *               boundary ((v + (q > 0)) - p0) / q (x, y, z on ties) until past
 *               the grid exit or the edge; at most 3 R + 3 steps
 */
#include <math.h>
#include <stdint.h>

void nlo_downsample(const uint32_t* child, uint32_t Rc, uint32_t* parent) {
    const uint32_t Rp = Rc / 2;
    const uint64_t n = (uint64_t)Rp * Rp * Rp;
    for (uint64_t w = 0; w < (n + 31) / 32; ++w) parent[w] = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t x = (uint32_t)(i % Rp), y = (uint32_t)((i / Rp) % Rp), z = (uint32_t)(i / ((uint64_t)Rp * Rp));
        int occ = 0;
        for (int c = 0; c < 8; ++c) {
            uint64_t ci = (uint64_t)(2 * x + (c & 1)) + (uint64_t)(2 * y + ((c >> 1) & 1)) * Rc +
                          (uint64_t)(2 * z + (c >> 2)) * Rc * Rc;
            occ |= (child[ci >> 5] >> (31u - (uint32_t)(ci & 31u))) & 1u;
        }
        if (occ) parent[i >> 5] |= 1u << (31u - (uint32_t)(i & 31u));
    }
}

void nlo_march(const uint32_t* bits, uint32_t res, const float bmin[3], const float unit[3], uint32_t n,
               const float* orig, const float* dirs, const float* t_min_in, const uint32_t* prev_in, uint32_t* vox,
               float* t_near, float* t_far) {
    const float R = (float)res;
    for (uint32_t i = 0; i < n; ++i) {
        float p0[3], q[3];
        for (int a = 0; a < 3; ++a) {
            p0[a] = (orig[(uint64_t)i * 3 + a] - bmin[a]) / unit[a];
            q[a] = dirs[(uint64_t)i * 3 + a] / unit[a];
        }
        const float t_min = t_min_in ? t_min_in[i] : 0.0f;
        const uint32_t prev = prev_in ? prev_in[i] : 0xffffffffu;
        uint32_t found = 0xffffffffu;
        float hit_near = -1.0f, hit_far = -1.0f;
        float tn = t_min, tf = INFINITY;
        int miss = 0;
        for (int a = 0; a < 3; ++a) {
            if (q[a] == 0.0f) {
                if (!(p0[a] >= 0.0f && p0[a] <= R)) miss = 1;
                continue;
            }
            float t0 = (0.0f - p0[a]) / q[a], t1 = (R - p0[a]) / q[a];
            if (t1 < t0) { float tmp = t0; t0 = t1; t1 = tmp; }
            if (t0 > tn) tn = t0;
            if (t1 < tf) tf = t1;
        }
        if (!miss && tn <= tf) {
            int v[3];
            for (int a = 0; a < 3; ++a) {
                float c = floorf(p0[a] + tn * q[a]);
                v[a] = c < 0.0f ? 0 : (c > R - 1.0f ? (int)res - 1 : (int)c);
            }
            const int max_steps = 3 * (int)res + 3;
            for (int step = 0; step < max_steps; ++step) {
                uint32_t lin = (uint32_t)v[0] + (uint32_t)v[1] * res + (uint32_t)v[2] * res * res;
                if (lin != prev && ((bits[lin >> 5] >> (31u - (lin & 31u))) & 1u)) {
                    float bn = -INFINITY, bf = INFINITY;
                    for (int a = 0; a < 3; ++a) {
                        if (q[a] == 0.0f) continue;
                        float t0 = ((float)v[a] - p0[a]) / q[a], t1 = ((float)(v[a] + 1) - p0[a]) / q[a];
                        if (t1 < t0) { float tmp = t0; t0 = t1; t1 = tmp; }
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
                    float tb = ((float)(v[a] + (q[a] > 0.0f ? 1 : 0)) - p0[a]) / q[a];
                    if (tb < best) { best = tb; axis = a; }
                }
                if (axis < 0 || best > tf) break;
                v[axis] += q[axis] > 0.0f ? 1 : -1;
                if (v[axis] < 0 || v[axis] >= (int)res) break;
            }
        }
        vox[i] = found;
        t_near[i] = hit_near;
        t_far[i] = hit_far;
    }
}
