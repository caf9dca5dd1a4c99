// The neural LoD path tracer's per-frame loop on gfx950.
//
// WavefrontNeuralThroughputVisibilityLod::voxel_tracing
// (rptr/cuda/neural-lod/wavefront_neural_throughput_visibility_lod.cu:464-720)
// traces one sample per pixel through a LoD level of the occupancy pyramid:
// per bounce, up to 3 R rounds of march to the next occupied voxel -> compact
// -> visibility network -> threshold test, then one throughput-network query
// per path that stopped in a voxel, which scales the path throughput and
// turns the ray.  The reference reads the active-ray count back to the host
// after every round and every bounce (:551-556, :619-624, :705-710).
//
// Two schedules compute the same per-ray arithmetic (the device functions
// below), so their images agree bit for bit:
//   REFERENCE   the reference's wavefront order: stage kernels over ray
//               buffers, compaction with atomic counters and a host readback
//               of each count, the networks through nlod_model_inference.
//               Any precision; the plain FP32 form is the baseline.
//   PERSISTENT  FP16 models: one launch for the whole frame.  Each lane of a
//               wave owns one ray and runs its state machine (march ->
//               visibility -> ... -> throughput -> bounce); when the path
//               ends, the lane takes the next ray from a device queue.  Each
//               round the wave marches its rays, then the lanes waiting for
//               whichever network more of them wait for write their inputs
//               to LDS rows, and the whole wave evaluates the rows as 16-query
//               MFMA tiles (both networks' weights in LDS) and hands each
//               lane its outputs back through LDS.  No host readback, no
//               grid-wide barrier; a wave exits when the queue is empty and
//               its rays are done, so the launch is capturable in a hipGraph.
// Per ray everything depends only on the ray's own state (the networks give a
// query's output independently of the other queries of its batch), so the
// order in which rays are processed cannot change a result.
//
// Arithmetic (no FP contraction; division and square root correctly rounded):
// GLM vector operations component-wise, dot = (x x' + y y') + z z', normalize
// = v * (1 / sqrt(dot)); the reference's double literals promote as in C++;
// sin, cos and exp are this file's double-precision routines (device and host
// math libraries differ in the last bit).  The march is a 3D DDA in the
// level's index space (origin in voxel units, world direction), the nanovdb
// TreeMarcher + DDA of first_active_voxel_intersection
// (neural_lod_learning_common.cuh:784-832); each occupied voxel other than the
// previous one is tested with intersect_aabb_min_max (:567-605).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "fused_net.h"
#include "model.h"
#include "nlod.h"
#include "nlod_common.h"

namespace nlod {

struct LodGrid;
void lod_grid_tracer_view(const LodGrid& g, uint32_t lod, const uint32_t** bits, const float** thr, uint32_t* res,
                          float bmin[3], float bmax[3], float* voxel_size, uint32_t* grid_res);

namespace {

// path_state values (wavefront_neural_pt.h:25-28)
constexpr int32_t kInactive = -3, kEscaped = -2, kActive = -1, kHit = 0;

struct TracerArgs {
    const uint32_t* bits;  // level bits, voxel x + y R + z R^2 in bit 31 - i % 32
    const float* thr;      // dense R^3 per-voxel thresholds, or null (0.5)
    uint32_t R, lod, grid_res, max_vis, max_depth, n_rays, n_thr_in;
    float bmin[3], ext[3], voxel_size;
    // camera (ViewParams: cam_pos, cam_dir_top_left, cam_du, cam_dv) and framebuffer
    float cam_pos[3], cam_tl[3], cam_du[3], cam_dv[3];
    uint32_t width, height, sample_idx, sample_offset;
    float background[3], envmap[3];
    float dilation;  // min(inference_min_extent_dilation, voxel_extent_dilation_outward)
    float max_rr_prob, hg_g;
    uint32_t apply_rr, rr_start_bounce, apply_vis_rr, stochastic, use_hg, log_learning;
};

// One path (NeuralBounceQuery + NeuralBounceResult, wavefront_neural_pt.h:43-62).
struct TRay {
    float o[3], d[3];
    uint32_t rng;
    int32_t state;
    int32_t vox[3];
    float t_near, t_far, thr;
    float T[3], wi[3], pdf_inv, illum[3];
    uint32_t pixel, needed, bounce;
};

// ------------------------------------------------------------------ rng
__device__ __forceinline__ uint32_t murmur_mix(uint32_t h, uint32_t k) {
    k *= 0xcc9e2d51u;
    k = (k << 15) | (k >> 17);
    k *= 0x1b873593u;
    h ^= k;
    return ((h << 13) | (h >> 19)) * 5u + 0xe6546b64u;
}
__device__ __forceinline__ uint32_t murmur_finalize(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    return h ^ (h >> 16);
}
// get_lcg_rng(index, frame, linear) (rptr/rendering/pointsets/lcg_rng.glsl:28-34)
__device__ __forceinline__ uint32_t lcg_seed(uint32_t index, uint32_t frame, uint32_t linear) {
    return murmur_finalize(murmur_mix(murmur_mix(frame, linear), index));
}
// lcg_randomf (lcg_rng.glsl:17-26)
__device__ __forceinline__ float lcg_next(uint32_t& s) {
    s = s * 1664525u + 1013904223u;
    return ldexpf((float)s, -32);
}

// ------------------------------------------------------------------ math
// sin and cos, double precision: reduction by pi/2 (two-part constant whose
// high part has trailing zeros, so k * hi is exact), Taylor on [-pi/4, pi/4].
__device__ __forceinline__ void sin_cos(double x, double& s, double& c) {
    const double k = rint(x * 0.63661977236758134308);
    const double r = (x - k * 1.5707963267341256) - k * 6.077100506506192e-11;
    const double r2 = r * r;
    double ps = -1.0 / 1307674368000.0;
    ps = ps * r2 + 1.0 / 6227020800.0;
    ps = ps * r2 + -1.0 / 39916800.0;
    ps = ps * r2 + 1.0 / 362880.0;
    ps = ps * r2 + -1.0 / 5040.0;
    ps = ps * r2 + 1.0 / 120.0;
    ps = ps * r2 + -1.0 / 6.0;
    ps = r + (r * r2) * ps;
    double pc = -1.0 / 87178291200.0;
    pc = pc * r2 + 1.0 / 479001600.0;
    pc = pc * r2 + -1.0 / 3628800.0;
    pc = pc * r2 + 1.0 / 40320.0;
    pc = pc * r2 + -1.0 / 720.0;
    pc = pc * r2 + 1.0 / 24.0;
    pc = pc * r2 + -0.5;
    pc = 1.0 + r2 * pc;
    switch (((int)k) & 3) {
        case 0: s = ps; c = pc; break;
        case 1: s = pc; c = -ps; break;
        case 2: s = -ps; c = -pc; break;
        default: s = -pc; c = ps; break;
    }
}
// exp, double precision: x = k ln2 + r (two-part ln2), Taylor to r^13, times 2^k.
__device__ __forceinline__ double exp_d(double x) {
    if (x > 709.0) return INFINITY;
    if (x < -745.0) return 0.0;
    const double k = rint(x * 1.4426950408889634);
    const double r = (x - k * 0.6931471803691238) - k * 1.9082149292705877e-10;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, (int)k);
}

__device__ __forceinline__ float dot3(const float* a, const float* b) {
    const float t0 = a[0] * b[0], t1 = a[1] * b[1], t2 = a[2] * b[2];
    return (t0 + t1) + t2;
}

// square_to_uniform_sphere (rptr/rendering/util.glsl:224-230)
__device__ __forceinline__ void uniform_sphere(float u0, float u1, float* w) {
    const float phi = (float)(6.283185307179586 * (double)u0);
    const float cos_t = u1 * 2.0f - 1.0f;
    const float sin_t = sqrtf(1.0f - cos_t * cos_t);
    double sd, cd;
    sin_cos((double)phi, sd, cd);
    w[0] = sin_t * (float)cd;
    w[1] = sin_t * (float)sd;
    w[2] = cos_t;
}

// square_to_hg (util.glsl:232-262), w_o = the ray direction
__device__ __forceinline__ void sample_hg(float u0, float u1, const float* w_o, float g, float* w) {
    float cos_t;
    if ((double)fabsf(g) < 1e-3) {
        cos_t = (float)(1.0 - 2.0 * (double)u0);
    } else {
        const float gg = g * g;
        const float sq = (float)((1.0 - (double)gg) / ((1.0 - (double)g) + (2.0 * (double)g) * (double)u0));
        const float sq2 = sq * sq;
        cos_t = (float)(((1.0 + (double)gg) - (double)sq2) / (2.0 * (double)g));
    }
    const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    const float phi = (float)(6.283185307179586 * (double)u1);
    double sd, cd;
    sin_cos((double)phi, sd, cd);
    const float lx = sin_t * (float)cd, ly = sin_t * (float)sd, lz = cos_t;
    const float n[3] = {-w_o[0], -w_o[1], -w_o[2]};
    float s[3];
    if (fabsf(n[0]) > fabsf(n[1])) {
        const float il = 1.0f / sqrtf(n[0] * n[0] + n[2] * n[2]);
        s[0] = n[2] * il;
        s[1] = 0.0f;
        s[2] = -n[0] * il;
    } else {
        const float il = 1.0f / sqrtf(n[1] * n[1] + n[2] * n[2]);
        s[0] = 0.0f;
        s[1] = n[2] * il;
        s[2] = -n[1] * il;
    }
    const float t[3] = {n[1] * s[2] - s[1] * n[2], n[2] * s[0] - s[2] * n[0], n[0] * s[1] - s[0] * n[1]};
#pragma unroll
    for (int a = 0; a < 3; ++a) w[a] = (lx * s[a] + ly * t[a]) + lz * n[a];
}

// square_to_hg_pdf (util.glsl:264-267)
__device__ __forceinline__ float hg_pdf(float cos_wo_wi, float g) {
    const float gg = g * g;
    const float den = (float)((1.0 + (double)gg) + (2.0 * (double)g) * (double)cos_wo_wi);
    const float dd = den * sqrtf(den);
    return (float)(((0.25 * 0.31830988618379067154) * (1.0 - (double)gg)) / (double)dd);
}

// ------------------------------------------------------------------ stages
// bits: the level's occupancy words (global memory, or the block's LDS copy)
__device__ __forceinline__ bool occupied(const TracerArgs& a, const uint32_t* bits, int x, int y, int z) {
    const uint32_t i = (uint32_t)x + (uint32_t)y * a.R + (uint32_t)z * a.R * a.R;
    return (bits[i >> 5] >> (31u - (i & 31u))) & 1u;
}

// camera_raygen (rptr/cuda/wavefront_neural_pt.cu:9-49)
__device__ __forceinline__ void raygen(const TracerArgs& a, uint32_t i, TRay& r) {
    const uint32_t px = i % a.width, py = i / a.width;
    uint32_t rng = lcg_seed(a.sample_idx, a.sample_offset, i);
    const float xf = ((float)px + lcg_next(rng)) / (float)a.width;
    const float yf = ((float)py + lcg_next(rng)) / (float)a.height;
    float d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = (xf * a.cam_du[k] + yf * a.cam_dv[k]) + a.cam_tl[k];
    const float il = 1.0f / sqrtf(dot3(d, d));
    const float res = (float)(a.grid_res >> a.lod);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.d[k] = d[k] * il;
        r.o[k] = res * ((a.cam_pos[k] - a.bmin[k]) / a.ext[k]);
        r.vox[k] = -1;
        r.T[k] = 1.0f;
        r.wi[k] = 0.0f;
        r.illum[k] = 0.0f;
    }
    r.rng = rng;
    r.state = kActive;
    r.t_near = r.t_far = -1.0f;
    r.thr = 0.0f;
    r.pdf_inv = 0.0f;
    r.pixel = py * a.width + px;
    r.needed = 0;
    r.bounce = 0;
}

// raymarch_impl (wavefront_neural_pt.cuh:89-141): first occupied voxel other
// than the previous one from the ray's minimum time; false = escaped.  The
// occupancy words come from `bits` (a.bits or an LDS copy of them).
__device__ __forceinline__ bool march(const TracerArgs& a, const uint32_t* bits, TRay& r) {
    const int R = (int)a.R;
    const float Rf = (float)a.R;
    float inv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) inv[k] = 1.0f / r.d[k];
    const float t_min = (r.t_near != -1.0f && r.t_far != -1.0f) ? (r.t_near + r.t_far) * 0.5f : 0.0f;
    float tn = t_min, tf = INFINITY;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (r.d[k] == 0.0f) {
            if (!(r.o[k] >= 0.0f && r.o[k] <= Rf)) return false;
            continue;
        }
        float t0 = (0.0f - r.o[k]) * inv[k], t1 = (Rf - r.o[k]) * inv[k];
        if (t1 < t0) {
            const float tmp = t0;
            t0 = t1;
            t1 = tmp;
        }
        if (t0 > tn) tn = t0;
        if (t1 < tf) tf = t1;
    }
    if (!(tn <= tf)) return false;
    int v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float c = floorf(r.o[k] + tn * r.d[k]);
        v[k] = c < 0.0f ? 0 : (c > Rf - 1.0f ? R - 1 : (int)c);
    }
    const int max_steps = 3 * R + 3;
    for (int step = 0; step < max_steps; ++step) {
        if (occupied(a, bits, v[0], v[1], v[2]) && !(v[0] == r.vox[0] && v[1] == r.vox[1] && v[2] == r.vox[2])) {
            // intersect_aabb_min_max with the voxel's box
            float nearT = -1e20f, farT = 1e20f;
            bool hit = true;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float lo = (float)v[k], hi = lo + 1.0f;
                if (r.d[k] == 0.0f) {
                    if (r.o[k] < lo || r.o[k] > hi) hit = false;
                } else if (hit) {
                    float t1 = (lo - r.o[k]) * inv[k], t2 = (hi - r.o[k]) * inv[k];
                    if (t1 > t2) {
                        const float tmp = t1;
                        t1 = t2;
                        t2 = tmp;
                    }
                    nearT = fmaxf(t1, nearT);
                    farT = fminf(t2, farT);
                    if (!(nearT < farT) || farT < 0.0f) hit = false;
                }
            }
            if (hit) {
#pragma unroll
                for (int k = 0; k < 3; ++k) r.vox[k] = v[k];
                r.t_near = nearT;
                r.t_far = farT;
                const uint32_t i = (uint32_t)v[0] + (uint32_t)v[1] * a.R + (uint32_t)v[2] * a.R * a.R;
                r.thr = a.thr ? a.thr[i] : 0.5f;
                return true;
            }
        }
        // advance the axis whose next boundary is nearest (x, y, z on ties);
        // written with selects, without dynamically indexed register arrays
        int axis = -1, dv = 0;
        float best = INFINITY;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (r.d[k] == 0.0f) continue;
            const bool pos = r.d[k] > 0.0f;
            const float tb = ((float)(v[k] + (pos ? 1 : 0)) - r.o[k]) * inv[k];
            if (tb < best) {
                best = tb;
                axis = k;
                dv = pos ? 1 : -1;
            }
        }
        if (axis < 0 || best > tf) return false;
        bool out = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (axis == k) {
                v[k] += dv;
                out = v[k] < 0 || v[k] >= R;
            }
        }
        if (out) return false;
    }
    return false;
}

// compute_lod_voxel (neural_lod_learning_grid.cuh:50-76)
__device__ __forceinline__ void voxel_center(const TracerArgs& a, const int* vox, float* c, float& extent) {
    const float step = ldexpf(a.voxel_size, (int)a.lod);
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = (a.bmin[k] + (float)vox[k] * step) + step * 0.5f;
    extent = step;
}

// generate_visibility_inference_data_impl (wavefront_neural_pt.cuh:143-197)
__device__ __forceinline__ void visibility_inputs(const TracerArgs& a, const TRay& r, float* in) {
    float c[3], extent;
    voxel_center(a, r.vox, c, extent);
    float lp[2][3];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const float t = j == 0 ? r.t_near : r.t_far;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float pos = r.o[k] + r.d[k] * t;
            const float l = (pos - (float)r.vox[k]) - 0.5f;  // index_to_local_voxel_pos
            lp[j][k] = fminf(fmaxf(l, -0.5f), 0.5f);
        }
    }
    // get_deterministic_local_pi_po (neural_lod_learning_common.cuh:255-305)
    int face[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        float best = -2e20f;
        face[j] = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float dist = fabsf(lp[j][k]);
            if (dist > best) {
                face[j] = lp[j][k] > 0.0f ? 2 * k : 2 * k + 1;
                best = dist;
            }
        }
    }
    const bool swap = face[0] > face[1];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float src = swap ? lp[1 - j][k] : lp[j][k];
            const float l = src * (1.0f + a.dilation);
            const float w = c[k] + l * extent;
            in[3 * j + k] = (w - a.bmin[k]) / a.ext[k];
        }
    }
}

// generate_throughput_inference_data_impl (wavefront_neural_pt.cuh:199-254)
__device__ __forceinline__ void throughput_inputs(const TracerArgs& a, TRay& r, float* in) {
    const float u0 = lcg_next(r.rng);
    const float u1 = lcg_next(r.rng);
    float wi[3];
    if (a.use_hg) sample_hg(u0, u1, r.d, a.hg_g, wi);
    else uniform_sphere(u0, u1, wi);
    float c[3], extent;
    voxel_center(a, r.vox, c, extent);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        in[k] = (c[k] - a.bmin[k]) / a.ext[k];
        in[3 + k] = (-r.d[k] + 1.0f) * 0.5f;
        in[6 + k] = (wi[k] + 1.0f) * 0.5f;
        r.wi[k] = wi[k];
    }
    if (a.n_thr_in > 9) in[9] = (float)a.lod;
    r.pdf_inv = a.use_hg ? 1.0f / hg_pdf(dot3(r.d, wi), a.hg_g) : (float)(4.0f * 3.14159265358979323846);
}

// check_visibility (wavefront_neural_throughput_visibility_lod.cu:290-338),
// visibility_rr_terminate (wavefront_neural_pt.cuh:75-87)
__device__ __forceinline__ void check_visibility(const TracerArgs& a, TRay& r, float v, uint32_t k) {
    r.needed += 1;
    const bool occluded = a.stochastic ? v >= lcg_next(r.rng) : v >= r.thr;
    if (occluded) {
        r.state = kHit;
    } else if (a.apply_vis_rr && k >= 5) {
        const float prob = fmaxf(v, a.max_rr_prob);
        if (lcg_next(r.rng) < prob) {
#pragma unroll
            for (int c = 0; c < 3; ++c) r.T[c] = r.T[c] / prob;
        } else {
            r.state = kInactive;
        }
    }
}

// process_ray_bounce (:202-254) with get_throughput_inference
// (wavefront_neural_pt.cuh:256-273) and rr_terminate (:57-73)
__device__ __forceinline__ void bounce(const TracerArgs& a, TRay& r, const float* out, uint32_t b) {
    float th[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float v = out[c];
        if (a.log_learning) v = (float)exp_d((double)v) - 1.0f;
        th[c] = fmaxf(0.0f, v);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) r.T[c] = r.T[c] * (th[c] * r.pdf_inv);
    bool cont = r.T[0] != 0.0f || r.T[1] != 0.0f || r.T[2] != 0.0f;
    const float t_eps = (float)((double)r.t_near + 1e-4 * (double)(r.t_far - r.t_near));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r.o[c] = r.o[c] + r.d[c] * t_eps;
        r.d[c] = r.wi[c];
    }
    r.t_near = r.t_far = -1.0f;
    if (a.apply_rr && cont && b >= a.rr_start_bounce) {
        const float prob = (float)fmin(1.0, (double)fmaxf(r.T[0], fmaxf(r.T[1], r.T[2])));
        if (lcg_next(r.rng) < prob) {
#pragma unroll
            for (int c = 0; c < 3; ++c) r.T[c] = r.T[c] / prob;
        } else {
            cont = false;
        }
    }
    r.state = cont ? kActive : kInactive;
}

// process_final_bounce (:37-82, constant environment) + accumulate_neural_bounce
__device__ __forceinline__ void finish(const TracerArgs& a, TRay& r, uint32_t b, float* illum, uint32_t* inferences,
                       uint32_t* bounces, bool write) {
    if (r.state == kEscaped) {
#pragma unroll
        for (int c = 0; c < 3; ++c) r.illum[c] = b == 0 ? a.background[c] : r.T[c] * a.envmap[c];
    }
    r.state = kInactive;
    if (!write) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) illum[(size_t)r.pixel * 3 + c] = r.illum[c];
    if (inferences) inferences[r.pixel] = r.needed;
    if (bounces) bounces[r.pixel] = b;
}

// ------------------------------------------------------ reference schedule
constexpr int kStageThreads = 256;

// counters[0]: rays compacted for the next round; counters[1]: rays that
// stopped in a voxel during the current bounce
__global__ void zero_counters_kernel(uint32_t* c, uint32_t n) {
    if (threadIdx.x < n) c[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kStageThreads) void raygen_kernel(TracerArgs a, TRay* rays) {
    const uint32_t i = blockIdx.x * kStageThreads + threadIdx.x;
    if (i < a.n_rays) raygen(a, i, rays[i]);
}

__global__ __launch_bounds__(kStageThreads) void march_kernel(TracerArgs a, uint32_t n, TRay* rays) {
    const uint32_t i = blockIdx.x * kStageThreads + threadIdx.x;
    if (i >= n) return;
    TRay r = rays[i];
    if (r.state != kActive) return;
    if (!march(a, a.bits, r)) r.state = kEscaped;
    rays[i] = r;
}

// compact_for_visibility_and_accumulate (:84-137); counters[0] active, [1] hit
__global__ __launch_bounds__(kStageThreads) void compact_visibility_kernel(
    TracerArgs a, uint32_t n, const TRay* src, TRay* dst, TRay* hit, uint32_t* counters, int last_inference,
    int last_bounce, uint32_t b, float* illum, uint32_t* inferences, uint32_t* bounces) {
    const uint32_t i = blockIdx.x * kStageThreads + threadIdx.x;
    if (i >= n) return;
    TRay r = src[i];
    if (r.state == kActive && !last_inference) {
        dst[atomicAdd(&counters[0], 1u)] = r;
    } else if (r.state == kHit && !last_bounce) {
        hit[atomicAdd(&counters[1], 1u)] = r;
    } else {
        finish(a, r, b, illum, inferences, bounces, true);
    }
}

__global__ __launch_bounds__(kStageThreads) void visibility_inputs_kernel(TracerArgs a, uint32_t n,
                                                                          const TRay* rays, float* in) {
    const uint32_t i = blockIdx.x * kStageThreads + threadIdx.x;
    if (i >= n) return;
    const TRay r = rays[i];
    float x[6];
    visibility_inputs(a, r, x);
#pragma unroll
    for (int k = 0; k < 6; ++k) in[(size_t)i * 6 + k] = x[k];
}

__global__ __launch_bounds__(kStageThreads) void check_kernel(TracerArgs a, uint32_t n, TRay* rays,
                                                              const float* out, uint32_t n_out, uint32_t k) {
    const uint32_t i = blockIdx.x * kStageThreads + threadIdx.x;
    if (i >= n) return;
    TRay r = rays[i];
    if (r.state != kActive) return;
    check_visibility(a, r, out[(size_t)i * n_out], k);
    rays[i] = r;
}

__global__ __launch_bounds__(kStageThreads) void throughput_inputs_kernel(TracerArgs a, uint32_t n, TRay* rays,
                                                                          float* in) {
    const uint32_t i = blockIdx.x * kStageThreads + threadIdx.x;
    if (i >= n) return;
    TRay r = rays[i];
    float x[10];
    throughput_inputs(a, r, x);
    for (uint32_t k = 0; k < a.n_thr_in; ++k) in[(size_t)i * a.n_thr_in + k] = x[k];
    rays[i] = r;
}

// process_ray_bounce, then compact_for_next_bounce_and_accumulate (:139-182)
__global__ __launch_bounds__(kStageThreads) void bounce_compact_kernel(TracerArgs a, uint32_t n, const TRay* hit,
                                                                       const float* out, uint32_t n_out, uint32_t b,
                                                                       TRay* dst, uint32_t* counters, float* illum,
                                                                       uint32_t* inferences, uint32_t* bounces) {
    const uint32_t i = blockIdx.x * kStageThreads + threadIdx.x;
    if (i >= n) return;
    TRay r = hit[i];
    const float o3[3] = {out[(size_t)i * n_out], out[(size_t)i * n_out + 1], out[(size_t)i * n_out + 2]};
    bounce(a, r, o3, b);
    if (r.state == kActive) dst[atomicAdd(&counters[0], 1u)] = r;
    else finish(a, r, b, illum, inferences, bounces, true);
}

// ------------------------------------------------------ persistent schedule
constexpr int kTraceBlock = 768;
constexpr int kTraceWaves = kTraceBlock / 64;
constexpr int kSlots = 64;     // rays in flight per wave: one per lane
constexpr int kInStride = 10;  // floats of one slot's network inputs in LDS (<= 10 used)
constexpr int kOutStride = 3;  // floats of one slot's network outputs in LDS (<= 3 used)
constexpr size_t kMaxTraceLds = 160 * 1024;  // LDS of one CU
enum : int { kWantMarch = 0, kWantVis = 1, kWantThr = 2 };

// number of set bits of mask below this lane
__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// LB: the level's occupancy words are copied into LDS after the weights and
// the input rows (levels up to about 80^3), so the march's dependent bit
// reads cost an LDS round trip instead of an L2 one.
template <int W, int F, int MAXD, bool LB>
__global__ __launch_bounds__(kTraceBlock, 3) void trace_persistent_kernel(TracerArgs a, NetArgs vis, NetArgs thr,
                                                                       const half_t* __restrict__ packed_vis,
                                                                       const half_t* __restrict__ packed_thr,
                                                                       uint32_t* __restrict__ queue,
                                                                       float* __restrict__ illum,
                                                                       uint32_t* __restrict__ inferences,
                                                                       uint32_t* __restrict__ bounces) {
    extern __shared__ float4 smem4[];
    half_t* sw = reinterpret_cast<half_t*>(smem4);
    float* s_in = reinterpret_cast<float*>(sw + vis.packed_halves + thr.packed_halves);
    float* s_out = s_in + kTraceWaves * kSlots * kInStride;
    for (uint32_t k = threadIdx.x; k < vis.packed_halves / 8; k += kTraceBlock)
        reinterpret_cast<float4*>(sw + vis.lds_offset)[k] = reinterpret_cast<const float4*>(packed_vis)[k];
    for (uint32_t k = threadIdx.x; k < thr.packed_halves / 8; k += kTraceBlock)
        reinterpret_cast<float4*>(sw + thr.lds_offset)[k] = reinterpret_cast<const float4*>(packed_thr)[k];
    const uint32_t* bits = a.bits;
    if constexpr (LB) {
        uint32_t* s_bits = reinterpret_cast<uint32_t*>(s_out + kTraceWaves * kSlots * kOutStride);
        const uint32_t words = (a.R * a.R * a.R + 31u) / 32u;
        for (uint32_t k = threadIdx.x; k < words; k += kTraceBlock) s_bits[k] = a.bits[k];
        bits = s_bits;
    }
    __syncthreads();  // the last block-wide barrier: from here on waves run independently

    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15, wave = threadIdx.x >> 6;
    // the wave's query rows (network inputs) and output rows, one per slot
    float* rows = s_in + (size_t)wave * kSlots * kInStride;
    float* outs = s_out + (size_t)wave * kSlots * kOutStride;
    TRay r;
    bool live = false;
    int want = kWantMarch;
    uint32_t k = 0, b = 0;        // visibility round and bounce of the lane's ray
    uint32_t next = 0, end = 0;   // rays taken from the queue and not yet started
    bool drained = false;
    // wave-uniform tallies for the frame's statistics: network evaluations
    // (16-query tiles) and the queries they served, per network, and rounds
    uint32_t n_vis_tiles = 0, n_vis_q = 0, n_thr_tiles = 0, n_thr_q = 0, n_rounds = 0;
    for (;;) {
        // refill free lanes in order from the wave's block of the queue
        uint64_t free_lanes = __ballot(!live);
        while (free_lanes && !drained) {
            if (next == end) {
                uint32_t start = 0;
                if (lane == 0) start = atomicAdd(queue, (uint32_t)kSlots);
                start = __shfl(start, 0, 64);
                if (start >= a.n_rays) {
                    drained = true;
                    break;
                }
                next = start;
                end = min(start + (uint32_t)kSlots, a.n_rays);
            }
            const uint32_t take = min((uint32_t)__popcll(free_lanes), end - next);
            const uint32_t rank = lane_rank(free_lanes);
            if (!live && rank < take) {
                raygen(a, next + rank, r);
                live = true;
                want = kWantMarch;
                k = 0;
                b = 0;
            }
            next += take;
            free_lanes = __ballot(!live);
        }
        if (__ballot(live) == 0) break;  // the queue is empty and every lane is done
        ++n_rounds;

        if (live && want == kWantMarch) {
            if (!march(a, bits, r)) {
                r.state = kEscaped;
                finish(a, r, b, illum, inferences, bounces, true);
                live = false;
            } else if (k == a.max_vis - 1) {  // last round of the bounce: the path ends unresolved
                finish(a, r, b, illum, inferences, bounces, true);
                live = false;
            } else {
                want = kWantVis;
            }
        }
        const uint32_t n_vis = (uint32_t)__popcll(__ballot(live && want == kWantVis));
        const uint32_t n_thr = (uint32_t)__popcll(__ballot(live && want == kWantThr));
        if (n_vis == 0 && n_thr == 0) continue;
        // every lane waiting for the network that more lanes wait for writes
        // its inputs to the row of its rank; the wave then evaluates the rows
        // 16 at a time (lanes of query n of a tile read row 16 t + n; rows past
        // the count read row 0, whose output is not used)
        const bool use_thr = n_thr >= n_vis;
        const bool mine = live && want == (use_thr ? kWantThr : kWantVis);
        const uint64_t chosen = __ballot(mine);
        const uint32_t cnt = (uint32_t)__popcll(chosen);
        const uint32_t my_row = lane_rank(chosen);
        if (mine) {
            float in[10];
            if (use_thr) throughput_inputs(a, r, in);
            else visibility_inputs(a, r, in);
            const uint32_t n_in = use_thr ? thr.n_in : 6u;
#pragma unroll
            for (int j = 0; j < 10; ++j)
                if ((uint32_t)j < n_in) rows[my_row * kInStride + j] = in[j];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t tiles = (cnt + 15) / 16;
        for (uint32_t t = 0; t < tiles; ++t) {
            const uint32_t q = 16 * t + (uint32_t)n;
            const float* x = rows + (q < cnt ? q : 0u) * kInStride;
            if (use_thr) {
                const floatx4 o = run_net<W, F, MAXD>(thr, sw, x);
                if (g == 0 && q < cnt) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) outs[q * kOutStride + c] = act_fwd(thr.out_act, o[c]);
                }
            } else {
                const floatx4 o = run_net<W, F, MAXD>(vis, sw, x);
                if (g == 0 && q < cnt) outs[q * kOutStride] = act_fwd(vis.out_act, o[0]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (use_thr) {
            n_thr_tiles += tiles;
            n_thr_q += cnt;
        } else {
            n_vis_tiles += tiles;
            n_vis_q += cnt;
        }
        if (!mine) continue;
        if (use_thr) {
            float out[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) out[c] = outs[my_row * kOutStride + c];
            bounce(a, r, out, b);
            if (r.state == kActive) {
                ++b;
                k = 0;
                want = kWantMarch;
            } else {
                finish(a, r, b, illum, inferences, bounces, true);
                live = false;
            }
        } else {
            check_visibility(a, r, outs[my_row * kOutStride], k);
            if (r.state == kHit) {
                if (b == a.max_depth - 1) {  // the last bounce keeps no intersection
                    finish(a, r, b, illum, inferences, bounces, true);
                    live = false;
                } else {
                    want = kWantThr;
                }
            } else if (r.state == kInactive) {
                finish(a, r, b, illum, inferences, bounces, true);
                live = false;
            } else {
                ++k;
                want = kWantMarch;
            }
        }
    }
    if (lane == 0) {
        unsigned long long* st = reinterpret_cast<unsigned long long*>(queue + 16);
        atomicAdd(st + 0, (unsigned long long)n_vis_tiles);
        atomicAdd(st + 1, (unsigned long long)n_vis_q);
        atomicAdd(st + 2, (unsigned long long)n_thr_tiles);
        atomicAdd(st + 3, (unsigned long long)n_thr_q);
        atomicAdd(st + 4, (unsigned long long)n_rounds);
    }
}

// queue[0]: next ray; queue[16 ..]: the frame's statistics (uint64 each)
__global__ void reset_queue_kernel(uint32_t* q) {
    if (threadIdx.x < 64) q[threadIdx.x] = 0;
}

}  // namespace

struct Tracer {
    const LodGrid& grid;
    Model& vis;
    Model& thr;
    bool persistent_ok = false;
    uint32_t F = 2, maxdeg = 4;
    size_t smem = 0;
    DevBuf rays_a, rays_b, rays_hit, vis_in, vis_out, thr_in, thr_out, counters, queue;
    hipEvent_t ev = nullptr;
    uint32_t* host_counts = nullptr;  // pinned readback of the reference schedule's counters

    Tracer(const LodGrid& g, Model& v, Model& t) : grid(g), vis(v), thr(t) {
        if (v.precision != t.precision) fail(NLOD_ERR_INVALID_ARGUMENT, "both networks must use the same precision");
        if (v.L.n_input_dims != 6)
            fail(NLOD_ERR_INVALID_ARGUMENT, "the visibility network must take 6 inputs (entry and exit points)");
        if (v.L.net.n_output < 1) fail(NLOD_ERR_INVALID_ARGUMENT, "the visibility network has no output");
        if (t.L.n_input_dims != 9 && t.L.n_input_dims != 10)
            fail(NLOD_ERR_INVALID_ARGUMENT,
                 "the throughput network must take 9 inputs (voxel, wo, wi) or 10 (with the LoD level)");
        if (t.L.net.n_output < 3) fail(NLOD_ERR_INVALID_ARGUMENT, "the throughput network needs 3 outputs");
        const size_t in_bytes = (size_t)kTraceWaves * kSlots * (kInStride + kOutStride) * sizeof(float);
        persistent_ok = fused_pair_ok(v, t, kMaxTraceLds - in_bytes);
        if (persistent_ok) {
            uint32_t Fv, Ft, dv, dt;
            layout_features(v.L, Fv, dv);
            layout_features(t.L, Ft, dt);
            F = Fv;
            maxdeg = std::max(dv, dt);
            smem = mfma_packed_bytes(v.L.net) + mfma_packed_bytes(t.L.net) + in_bytes;
        }
        NLOD_HIP(hipHostMalloc((void**)&host_counts, 4 * sizeof(uint32_t), hipHostMallocDefault));
        NLOD_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    ~Tracer() {
        if (host_counts) (void)hipHostFree(host_counts);
        if (ev) (void)hipEventDestroy(ev);
    }

    TracerArgs args(const nlod_camera& cam, const nlod_tracer_params& p, uint32_t sample_idx,
                    uint32_t sample_offset) const {
        TracerArgs a{};
        float bmax[3];
        lod_grid_tracer_view(grid, p.lod, &a.bits, &a.thr, &a.R, a.bmin, bmax, &a.voxel_size, &a.grid_res);
        a.lod = p.lod;
        a.max_vis = (a.grid_res >> p.lod) * 3;  // Manhattan distance across the level (:489)
        a.max_depth = p.max_depth;
        a.n_rays = cam.width * cam.height;
        a.n_thr_in = thr.L.n_input_dims;
        for (int k = 0; k < 3; ++k) {
            a.ext[k] = bmax[k] - a.bmin[k];
            a.cam_pos[k] = cam.position[k];
            a.cam_tl[k] = cam.dir_top_left[k];
            a.cam_du[k] = cam.du[k];
            a.cam_dv[k] = cam.dv[k];
            a.background[k] = p.background_color[k];
            a.envmap[k] = p.envmap_color[k];
        }
        a.width = cam.width;
        a.height = cam.height;
        a.sample_idx = sample_idx;
        a.sample_offset = sample_offset;
        a.dilation = std::min(p.inference_min_extent_dilation, p.voxel_extent_dilation_outward);
        // max_visibility_rr_prob_const (wavefront_neural_pt.cuh:347)
        a.max_rr_prob = powf(1e-6f, 1.f / (float)(3 * a.grid_res));
        a.hg_g = p.hg_g;
        a.apply_rr = p.apply_rr;
        a.rr_start_bounce = p.rr_start_bounce;
        a.apply_vis_rr = p.apply_visibility_rr;
        a.stochastic = p.stochastic_threshold;
        a.use_hg = p.use_hg;
        a.log_learning = p.throughput_log_learning;
        return a;
    }

    void check_args(const nlod_camera& cam, const nlod_tracer_params& p) const {
        NLOD_CHECK(cam.width > 0 && cam.height > 0, "empty framebuffer");
        NLOD_CHECK((uint64_t)cam.width * cam.height <= (1ull << 28), "framebuffer above 2^28 pixels");
        NLOD_CHECK(p.max_depth >= 1, "max_depth must be at least 1");
        NLOD_CHECK(p.schedule <= NLOD_TRACER_PERSISTENT, "unknown schedule");
        NLOD_CHECK(std::isfinite(p.hg_g) && std::fabs(p.hg_g) < 1.0f, "hg_g must lie in (-1, 1)");
    }

    int pick(const nlod_tracer_params& p) const {
        if (p.schedule == NLOD_TRACER_REFERENCE) return NLOD_TRACER_REFERENCE;
        if (p.schedule == NLOD_TRACER_PERSISTENT) {
            if (!persistent_ok)
                fail(NLOD_ERR_UNSUPPORTED,
                     "the persistent schedule needs two FP16 networks of width 32 or 64 with one feature count "
                     "whose weights fit in LDS");
            return NLOD_TRACER_PERSISTENT;
        }
        return persistent_ok ? NLOD_TRACER_PERSISTENT : NLOD_TRACER_REFERENCE;
    }

    void render(hipStream_t s, const nlod_camera& cam, const nlod_tracer_params& p, uint32_t sample_idx,
                uint32_t sample_offset, float* illum, uint32_t* inferences, uint32_t* bounces) {
        check_args(cam, p);
        vis.require_params(true);
        thr.require_params(true);
        const TracerArgs a = args(cam, p, sample_idx, sample_offset);
        if (pick(p) == NLOD_TRACER_PERSISTENT) render_persistent(s, a, illum, inferences, bounces);
        else render_reference(s, a, illum, inferences, bounces);
    }

    void render_persistent(hipStream_t s, const TracerArgs& a, float* illum, uint32_t* inferences,
                           uint32_t* bounces) {
        queue.ensure(sizeof(uint32_t) * 64);
        const NetArgs av = net_args_of(vis, 0);
        const NetArgs at = net_args_of(thr, av.packed_halves);
        hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, s, queue.as<uint32_t>());
        NLOD_LAUNCH_CHECK();
        // the level's occupancy words join the weights in LDS when they fit
        const size_t bits_bytes = sizeof(uint32_t) * (((size_t)a.R * a.R * a.R + 31) / 32);
        const bool lds_bits = smem + bits_bytes <= kMaxTraceLds;
        const size_t launch_smem = lds_bits ? smem + bits_bytes : smem;
        dispatch_wf(vis.L.net.width, F, maxdeg, [&](auto Wc, auto Fc, auto Dc) {
            constexpr int Wv = decltype(Wc)::value, Fv = decltype(Fc)::value, Dv = decltype(Dc)::value;
            auto kern = lds_bits ? trace_persistent_kernel<Wv, Fv, Dv, true> : trace_persistent_kernel<Wv, Fv, Dv, false>;
            const int per_cu = std::max(1, occupancy_blocks((const void*)kern, kTraceBlock, launch_smem));
            const uint32_t waves_needed = div_round_up(a.n_rays, 16);
            const uint32_t blocks = std::max<uint32_t>(
                1, std::min<uint32_t>(div_round_up(waves_needed, kTraceWaves), (uint32_t)(cu_count() * per_cu)));
            hipLaunchKernelGGL(kern, dim3(blocks), dim3(kTraceBlock), launch_smem, s, a, av, at,
                               static_cast<const half_t*>(vis.packed_fwd(true)),
                               static_cast<const half_t*>(thr.packed_fwd(true)), queue.as<uint32_t>(), illum,
                               inferences, bounces);
        });
        NLOD_LAUNCH_CHECK();
    }

    // statistics of the last persistent frame (synchronizes the stream)
    void stats(hipStream_t s, uint64_t* out) {
        std::memset(out, 0, 5 * sizeof(uint64_t));
        if (!queue.p) return;
        NLOD_HIP(hipMemcpyAsync(out, queue.as<uint32_t>() + 16, 5 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        NLOD_HIP(hipStreamSynchronize(s));
    }

    uint32_t read_count(hipStream_t s, int which) {
        NLOD_HIP(hipMemcpyAsync(host_counts + which, counters.as<uint32_t>() + which, sizeof(uint32_t),
                                hipMemcpyDeviceToHost, s));
        NLOD_HIP(hipStreamSynchronize(s));
        return host_counts[which];
    }

    // voxel_tracing in the reference's order, one count readback per round
    void render_reference(hipStream_t s, const TracerArgs& a, float* illum, uint32_t* inferences,
                          uint32_t* bounces) {
        const uint32_t n = a.n_rays;
        rays_a.ensure(sizeof(TRay) * n);
        rays_b.ensure(sizeof(TRay) * n);
        rays_hit.ensure(sizeof(TRay) * n);
        vis_in.ensure(sizeof(float) * 6 * (size_t)n);
        vis_out.ensure(sizeof(float) * vis.L.net.n_output * (size_t)n);
        thr_in.ensure(sizeof(float) * a.n_thr_in * (size_t)n);
        thr_out.ensure(sizeof(float) * thr.L.net.n_output * (size_t)n);
        counters.ensure(sizeof(uint32_t) * 4);
        TRay* cur = rays_a.as<TRay>();
        TRay* nxt = rays_b.as<TRay>();
        TRay* hit = rays_hit.as<TRay>();
        uint32_t* cnt = counters.as<uint32_t>();
        auto grid_of = [](uint32_t m) { return dim3(div_round_up(m, kStageThreads)); };
        hipLaunchKernelGGL(raygen_kernel, grid_of(n), dim3(kStageThreads), 0, s, a, cur);
        NLOD_LAUNCH_CHECK();
        uint32_t n_active = n;
        for (uint32_t b = 0; b < a.max_depth && n_active > 0; ++b) {
            hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, cnt, 2u);
            NLOD_LAUNCH_CHECK();
            for (uint32_t k = 0; k < a.max_vis; ++k) {
                hipLaunchKernelGGL(march_kernel, grid_of(n_active), dim3(kStageThreads), 0, s, a, n_active, cur);
                NLOD_LAUNCH_CHECK();
                hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, cnt, 1u);
                NLOD_LAUNCH_CHECK();
                hipLaunchKernelGGL(compact_visibility_kernel, grid_of(n_active), dim3(kStageThreads), 0, s, a,
                                   n_active, cur, nxt, hit, cnt, (int)(k == a.max_vis - 1),
                                   (int)(b == a.max_depth - 1), b, illum, inferences, bounces);
                NLOD_LAUNCH_CHECK();
                n_active = read_count(s, 0);
                std::swap(cur, nxt);
                if (n_active == 0) break;
                hipLaunchKernelGGL(visibility_inputs_kernel, grid_of(n_active), dim3(kStageThreads), 0, s, a,
                                   n_active, cur, vis_in.as<float>());
                NLOD_LAUNCH_CHECK();
                vis.inference(s, n_active, vis_in.as<float>(), vis_out.as<float>());
                hipLaunchKernelGGL(check_kernel, grid_of(n_active), dim3(kStageThreads), 0, s, a, n_active, cur,
                                   vis_out.as<float>(), vis.L.net.n_output, k);
                NLOD_LAUNCH_CHECK();
            }
            const uint32_t n_hit = read_count(s, 1);
            if (n_hit == 0) break;
            hipLaunchKernelGGL(throughput_inputs_kernel, grid_of(n_hit), dim3(kStageThreads), 0, s, a, n_hit, hit,
                               thr_in.as<float>());
            NLOD_LAUNCH_CHECK();
            thr.inference(s, n_hit, thr_in.as<float>(), thr_out.as<float>());
            hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, cnt, 1u);
            NLOD_LAUNCH_CHECK();
            hipLaunchKernelGGL(bounce_compact_kernel, grid_of(n_hit), dim3(kStageThreads), 0, s, a, n_hit, hit,
                               thr_out.as<float>(), thr.L.net.n_output, b, cur, cnt, illum, inferences, bounces);
            NLOD_LAUNCH_CHECK();
            n_active = read_count(s, 0);
        }
    }
};

}  // namespace nlod

// ============================================================== C ABI
struct nlod_model;
struct nlod_lod_grid;
struct nlod_tracer {
    std::unique_ptr<nlod::Tracer> t;
};

namespace nlod {
Model& model_of(nlod_model* m);
const LodGrid& lod_grid_of(const nlod_lod_grid* g);
}  // namespace nlod

extern "C" {

int nlod_tracer_default_params(nlod_tracer_params* p) {
    return nlod::guarded([&] {
        NLOD_CHECK(p, "null argument");
        std::memset(p, 0, sizeof(*p));
        // WavefrontNeuralThroughputVisibilityLod params defaults (wavefront_neural_throughput_visibility_lod.h:19-31)
        p->max_depth = 3;
        for (int k = 0; k < 3; ++k) p->background_color[k] = p->envmap_color[k] = 1.0f;
        p->inference_min_extent_dilation = 1e-4f;
        p->voxel_extent_dilation_outward = 1e-4f;
        p->rr_start_bounce = 1;
        p->hg_g = 0.5f;
        p->schedule = NLOD_TRACER_AUTO;
    });
}

int nlod_tracer_create(const nlod_lod_grid* grid, nlod_model* visibility, nlod_model* throughput,
                       nlod_tracer** out) {
    return nlod::guarded([&] {
        NLOD_CHECK(grid && visibility && throughput && out, "null argument");
        *out = nullptr;
        auto h = std::make_unique<nlod_tracer>();
        h->t = std::make_unique<nlod::Tracer>(nlod::lod_grid_of(grid), nlod::model_of(visibility),
                                              nlod::model_of(throughput));
        *out = h.release();
    });
}

int nlod_tracer_destroy(nlod_tracer* tracer) {
    return nlod::guarded([&] { delete tracer; });
}

int nlod_tracer_render(nlod_tracer* tracer, void* stream, const nlod_camera* camera, const nlod_tracer_params* params,
                       uint32_t sample_idx, uint32_t sample_offset, float* d_illumination, uint32_t* d_inferences,
                       uint32_t* d_bounces) {
    return nlod::guarded([&] {
        NLOD_CHECK(tracer && camera && params && d_illumination, "null argument");
        tracer->t->render(nlod::as_stream(stream), *camera, *params, sample_idx, sample_offset, d_illumination,
                          d_inferences, d_bounces);
    });
}

int nlod_tracer_stats(nlod_tracer* tracer, void* stream, uint64_t* stats_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(tracer && stats_out, "null argument");
        tracer->t->stats(nlod::as_stream(stream), stats_out);
    });
}

int nlod_tracer_schedule(const nlod_tracer* tracer, const nlod_tracer_params* params, int* schedule_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(tracer && params && schedule_out, "null argument");
        *schedule_out = tracer->t->pick(*params);
    });
}

}  // extern "C"
