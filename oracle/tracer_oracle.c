/*
 * ORACLE — test infrastructure only.  Nothing in the product links, loads or
 * calls this file; only tests/ use it, and only as the checker.
 *
 * CPU restatement of the neural LoD path tracer's wavefront loop,
 * WavefrontNeuralThroughputVisibilityLod::voxel_tracing
 * (rptr/cuda/neural-lod/wavefront_neural_throughput_visibility_lod.cu:464-720)
 * with its stage kernels, in the reference's order:
 *   camera_raygen                          wavefront_neural_pt.cu:9-49
 *   raymarch / raymarch_impl               ..._lod.cu:184-200, wavefront_neural_pt.cuh:89-141
 *   compact_for_visibility_and_accumulate  ..._lod.cu:84-137 (+ process_final_bounce :37-82)
 *   generate_visibility_inference_data     wavefront_neural_pt.cuh:143-197
 *   check_visibility                       ..._lod.cu:290-338
 *   generate_throughput_inference_data     wavefront_neural_pt.cuh:199-254
 *   process_ray_bounce                     ..._lod.cu:202-254
 *   compact_for_next_bounce_and_accumulate ..._lod.cu:139-182
 * The two networks are a callback that receives a compacted batch (the
 * reference's tcnn inference calls, :595 and :662), so a test can supply the
 * GPU's own network outputs and compare every pixel bit for bit, or an FP64
 * network.  Per ray, everything is a function of its own state: the order in
 * which the GPU processes rays cannot change any result.
 *
 * Numerical conventions (the library follows the same ones; built with
 * -ffp-contract=off, every operation a separately rounded IEEE operation):
 *   - GLM vector ops component-wise, dot = (x x' + y y') + z z', normalize
 *     = v * (1 / sqrt(dot(v, v))); the reference's double literals (1e-4 in
 *     the bounce offset, M_PI, the HG sampler's 1.0 / 2.0) promote as C++ does.
 *   - sin / cos / exp are the library's own double-precision routines
 *     (restated here as nlo_sincos / nlo_exp), since device and host math
 *     libraries differ in the last bit; checked against libm in the tests.
 *   - Index-space march: the nanovdb TreeMarcher + DDA of
 *     first_active_voxel_intersection (neural_lod_learning_common.cuh:784-832)
 *     as a 3D DDA over the level's bits from the ray's minimum time; each
 *     occupied voxel other than the previous one is tested with
 *     intersect_aabb_min_max (:567-605, reciprocal direction, strict
 *     near < far, far >= 0).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NLO_INACTIVE (-3)
#define NLO_ESCAPED (-2)
#define NLO_ACTIVE (-1)
#define NLO_HIT 0 /* ACTIVE_PATH_WITH_INTERSECTION */

typedef void (*nlo_net_fn)(void* user, int which, uint32_t n, const float* in, float* out);

typedef struct {
    float position[3], dir_top_left[3], du[3], dv[3];
    uint32_t width, height;
} nlo_camera;

typedef struct {
    uint32_t lod, max_depth;
    float background[3], envmap[3];
    float inference_min_extent_dilation, voxel_extent_dilation_outward;
    uint32_t apply_rr, rr_start_bounce, apply_visibility_rr, stochastic_threshold, use_hg;
    float hg_g;
    uint32_t log_learning;
    uint32_t n_thr_inputs; /* 9: [centre, wo, wi]; 10: LoD level appended */
    uint32_t grid_res;     /* level-0 resolution (VOXEL_GRID_MAX_RES) */
} nlo_tracer_params;

typedef struct {
    const uint32_t* bits; /* level bits, voxel x + y R + z R^2 in bit 31 - i % 32 */
    const float* thr;     /* dense R^3 per-voxel thresholds, or NULL (0.5 each) */
    uint32_t R;
    float bmin[3], bmax[3];
    float voxel_size; /* level-0 voxel edge */
} nlo_tracer_scene;

typedef struct {
    float o[3], d[3];
    uint32_t rng;
    int state;
    int vox[3];
    float t_near, t_far, thr;
    float T[3], wi[3], pdf_inv, illum[3];
    uint32_t pixel, needed, bounce;
} nlo_ray;

/* ---------------------------------------------------------------- rng */
static uint32_t murmur_mix(uint32_t h, uint32_t k) {
    k *= 0xcc9e2d51u;
    k = (k << 15) | (k >> 17);
    k *= 0x1b873593u;
    h ^= k;
    h = ((h << 13) | (h >> 19)) * 5u + 0xe6546b64u;
    return h;
}
static uint32_t murmur_fin(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
/* get_lcg_rng(index, frame, linear) (lcg_rng.glsl:28-34) */
uint32_t nlo_lcg_seed(uint32_t index, uint32_t frame, uint32_t linear) {
    return murmur_fin(murmur_mix(murmur_mix(frame, linear), index));
}
/* lcg_randomf (lcg_rng.glsl:17-26): ldexp(float(state), -32) after one LCG step */
float nlo_lcg_next(uint32_t* s) {
    *s = *s * 1664525u + 1013904223u;
    return ldexpf((float)*s, -32);
}

/* ---------------------------------------------------------------- math */
/* sin and cos of x (|x| < 64), double precision: Cody-Waite reduction by pi/2
 * (hi part with trailing zeros), Taylor polynomials on [-pi/4, pi/4]. */
void nlo_sincos(double x, double* s, double* c) {
    const double k = rint(x * 0.63661977236758134308);
    const double r = (x - k * 1.5707963267341256) - k * 6.077100506506192e-11;
    const double r2 = r * r;
    const double ps =
        r + r * r2 *
                (-1.0 / 6 +
                 r2 * (1.0 / 120 +
                       r2 * (-1.0 / 5040 +
                             r2 * (1.0 / 362880 +
                                   r2 * (-1.0 / 39916800 + r2 * (1.0 / 6227020800.0 + r2 * (-1.0 / 1307674368000.0)))))));
    const double pc =
        1.0 + r2 * (-0.5 + r2 * (1.0 / 24 +
                                 r2 * (-1.0 / 720 +
                                       r2 * (1.0 / 40320 +
                                             r2 * (-1.0 / 3628800 +
                                                   r2 * (1.0 / 479001600.0 + r2 * (-1.0 / 87178291200.0)))))));
    switch (((int)k) & 3) {
        case 0: *s = ps; *c = pc; break;
        case 1: *s = pc; *c = -ps; break;
        case 2: *s = -ps; *c = -pc; break;
        default: *s = -pc; *c = ps; break;
    }
}
/* exp(x), double precision: x = k ln2 + r, Taylor on |r| <= ln2 / 2, times 2^k. */
double nlo_exp(double x) {
    if (x > 709.0) return INFINITY;
    if (x < -745.0) return 0.0;
    const double k = rint(x * 1.4426950408889634);
    const double r = (x - k * 0.6931471803691238) - k * 1.9082149292705877e-10;
    double p = 1.0 / 6227020800.0;
    const double inv[12] = {1.0 / 479001600, 1.0 / 39916800, 1.0 / 3628800, 1.0 / 362880, 1.0 / 40320,
                            1.0 / 5040,      1.0 / 720,       1.0 / 120,     1.0 / 24,     1.0 / 6,
                            0.5,             1.0};
    for (int i = 0; i < 12; ++i) p = p * r + inv[i];
    p = p * r + 1.0;
    return ldexp(p, (int)k);
}

static float dot3(const float* a, const float* b) {
    const float t0 = a[0] * b[0], t1 = a[1] * b[1], t2 = a[2] * b[2];
    return (t0 + t1) + t2;
}

/* square_to_uniform_sphere (util.glsl:224-230): phi as float, then sin/cos */
static void uniform_sphere(float u0, float u1, float* out) {
    const float phi = (float)(6.283185307179586 * (double)u0); /* 2.0f * M_PI * rng.x */
    const float cos_theta = u1 * 2.0f - 1.0f;
    const float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    double sd, cd;
    nlo_sincos((double)phi, &sd, &cd);
    out[0] = sin_theta * (float)cd;
    out[1] = sin_theta * (float)sd;
    out[2] = cos_theta;
}

/* square_to_hg (util.glsl:232-262) with w_o = the ray direction */
static void sample_hg(float u0, float u1, const float* w_o, float g, float* out) {
    float cos_theta;
    if ((double)fabsf(g) < 1e-3) {
        cos_theta = (float)(1.0 - 2.0 * (double)u0);
    } else {
        /* float sqr_term = (1.0 - g * g) / (1.0 - g + 2.0 * g * rng.x); */
        const float gg = g * g;
        const float sqr_term = (float)((1.0 - (double)gg) / ((1.0 - (double)g) + (2.0 * (double)g) * (double)u0));
        const float sq = sqr_term * sqr_term;
        cos_theta = (float)(((1.0 + (double)gg) - (double)sq) / (2.0 * (double)g));
    }
    const float sin_theta = sqrtf(fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
    const float phi = (float)(6.283185307179586 * (double)u1);
    double sd, cd;
    nlo_sincos((double)phi, &sd, &cd);
    const float local[3] = {sin_theta * (float)cd, sin_theta * (float)sd, cos_theta};
    const float n[3] = {-w_o[0], -w_o[1], -w_o[2]};
    float s[3];
    if (fabsf(n[0]) > fabsf(n[1])) {
        const float inv_len = 1.0f / sqrtf(n[0] * n[0] + n[2] * n[2]);
        s[0] = n[2] * inv_len; s[1] = 0.0f; s[2] = -n[0] * inv_len;
    } else {
        const float inv_len = 1.0f / sqrtf(n[1] * n[1] + n[2] * n[2]);
        s[0] = 0.0f; s[1] = n[2] * inv_len; s[2] = -n[1] * inv_len;
    }
    /* t = cross(n, s) (GLM: x = a.y b.z - b.y a.z, ...) */
    const float t[3] = {n[1] * s[2] - s[1] * n[2], n[2] * s[0] - s[2] * n[0], n[0] * s[1] - s[0] * n[1]};
    for (int a = 0; a < 3; ++a) out[a] = (local[0] * s[a] + local[1] * t[a]) + local[2] * n[a];
}

/* square_to_hg_pdf (util.glsl:264-267) */
static float hg_pdf(float wo_dot_wi, float g) {
    /* float denom = 1.0 + g * g + 2.0 * g * wo_dot_wi; */
    const float gg = g * g;
    const float denom = (float)((1.0 + (double)gg) + (2.0 * (double)g) * (double)wo_dot_wi);
    const float dd = denom * sqrtf(denom);
    return (float)(((0.25 * 0.31830988618379067154) * (1.0 - (double)gg)) / (double)dd);
}

/* ---------------------------------------------------------------- march */
static int occupied(const nlo_tracer_scene* sc, int x, int y, int z) {
    const uint64_t R = sc->R;
    const uint64_t i = (uint64_t)x + (uint64_t)y * R + (uint64_t)z * R * R;
    return (int)((sc->bits[i >> 5] >> (31u - (uint32_t)(i & 31u))) & 1u);
}

/* intersect_aabb_min_max (neural_lod_learning_common.cuh:567-605) */
static int box_hit(const float* o, const float* d, const float* inv, const float* lo, const float* hi, float* tn,
                   float* tf) {
    float nearT = -1e20f, farT = 1e20f;
    for (int a = 0; a < 3; ++a) {
        if (d[a] == 0.0f) {
            if (o[a] < lo[a] || o[a] > hi[a]) return 0;
        } else {
            float t1 = (lo[a] - o[a]) * inv[a], t2 = (hi[a] - o[a]) * inv[a];
            if (t1 > t2) { const float tmp = t1; t1 = t2; t2 = tmp; }
            nearT = fmaxf(t1, nearT);
            farT = fminf(t2, farT);
            if (!(nearT < farT) || farT < 0.0f) return 0;
        }
    }
    *tn = nearT;
    *tf = farT;
    return 1;
}

/* raymarch_impl: first occupied voxel other than the previous one along the
 * index-space ray from t_min; returns 0 when the ray leaves the grid. */
static int march(const nlo_tracer_scene* sc, nlo_ray* r) {
    const int R = (int)sc->R;
    const float Rf = (float)sc->R;
    float inv[3];
    for (int a = 0; a < 3; ++a) inv[a] = 1.0f / r->d[a];
    const float t_min = (r->t_near != -1.0f && r->t_far != -1.0f) ? (r->t_near + r->t_far) * 0.5f : 0.0f;
    /* clip to the grid box [0, R]^3 */
    float tn = t_min, tf = INFINITY;
    for (int a = 0; a < 3; ++a) {
        if (r->d[a] == 0.0f) {
            if (!(r->o[a] >= 0.0f && r->o[a] <= Rf)) return 0;
            continue;
        }
        float t0 = (0.0f - r->o[a]) * inv[a], t1 = (Rf - r->o[a]) * inv[a];
        if (t1 < t0) { const float tmp = t0; t0 = t1; t1 = tmp; }
        if (t0 > tn) tn = t0;
        if (t1 < tf) tf = t1;
    }
    if (!(tn <= tf)) return 0;
    int v[3];
    for (int a = 0; a < 3; ++a) {
        const float c = floorf(r->o[a] + tn * r->d[a]);
        v[a] = c < 0.0f ? 0 : (c > Rf - 1.0f ? R - 1 : (int)c);
    }
    const int max_steps = 3 * R + 3;
    for (int step = 0; step < max_steps; ++step) {
        if (occupied(sc, v[0], v[1], v[2]) && !(v[0] == r->vox[0] && v[1] == r->vox[1] && v[2] == r->vox[2])) {
            const float lo[3] = {(float)v[0], (float)v[1], (float)v[2]};
            const float hi[3] = {lo[0] + 1.0f, lo[1] + 1.0f, lo[2] + 1.0f};
            float bn, bf;
            if (box_hit(r->o, r->d, inv, lo, hi, &bn, &bf)) {
                r->vox[0] = v[0]; r->vox[1] = v[1]; r->vox[2] = v[2];
                r->t_near = bn;
                r->t_far = bf;
                const uint64_t i = (uint64_t)v[0] + (uint64_t)v[1] * sc->R + (uint64_t)v[2] * sc->R * sc->R;
                r->thr = sc->thr ? sc->thr[i] : 0.5f;
                return 1;
            }
        }
        int axis = -1;
        float best = INFINITY;
        for (int a = 0; a < 3; ++a) {
            if (r->d[a] == 0.0f) continue;
            const float tb = ((float)(v[a] + (r->d[a] > 0.0f ? 1 : 0)) - r->o[a]) * inv[a];
            if (tb < best) { best = tb; axis = a; }
        }
        if (axis < 0 || best > tf) return 0;
        v[axis] += r->d[axis] > 0.0f ? 1 : -1;
        if (v[axis] < 0 || v[axis] >= R) return 0;
    }
    return 0;
}

/* ---------------------------------------------------------------- stages */
static void voxel_of(const nlo_tracer_scene* sc, const nlo_tracer_params* p, const int* vox, float* center,
                     float* extent) {
    /* compute_lod_voxel (neural_lod_learning_grid.cuh:50-76) */
    const float step = ldexpf(sc->voxel_size, (int)p->lod);
    for (int a = 0; a < 3; ++a) {
        const float vmin = sc->bmin[a] + (float)vox[a] * step;
        center[a] = vmin + step * 0.5f;
    }
    *extent = step;
}

static void raygen(const nlo_tracer_scene* sc, const nlo_camera* cam, const nlo_tracer_params* p, uint32_t sample_idx,
                   uint32_t sample_offset, uint32_t i, nlo_ray* r) {
    const uint32_t px = i % cam->width, py = i / cam->width;
    uint32_t rng = nlo_lcg_seed(sample_idx, sample_offset, i);
    const float xf = ((float)px + nlo_lcg_next(&rng)) / (float)cam->width;
    const float yf = ((float)py + nlo_lcg_next(&rng)) / (float)cam->height;
    float d[3];
    for (int a = 0; a < 3; ++a) d[a] = (xf * cam->du[a] + yf * cam->dv[a]) + cam->dir_top_left[a];
    const float inv_len = 1.0f / sqrtf(dot3(d, d));
    const float res = (float)(p->grid_res >> p->lod);
    memset(r, 0, sizeof(*r));
    for (int a = 0; a < 3; ++a) {
        r->d[a] = d[a] * inv_len;
        r->o[a] = res * ((cam->position[a] - sc->bmin[a]) / (sc->bmax[a] - sc->bmin[a]));
        r->vox[a] = -1;
        r->T[a] = 1.0f;
    }
    r->rng = rng;
    r->state = NLO_ACTIVE;
    r->t_near = r->t_far = -1.0f;
    r->pixel = py * cam->width + px; /* coord_2d_to_1d */
}

/* process_final_bounce (use_envmap = false) */
static void finalize(const nlo_tracer_params* p, nlo_ray* r, uint32_t bounce) {
    if (r->state == NLO_ESCAPED) {
        for (int a = 0; a < 3; ++a) r->illum[a] = bounce == 0 ? p->background[a] : r->T[a] * p->envmap[a];
    }
    r->state = NLO_INACTIVE;
    r->bounce = bounce;
}

static void visibility_inputs(const nlo_tracer_scene* sc, const nlo_tracer_params* p, const nlo_ray* r, float* in) {
    float center[3], extent;
    voxel_of(sc, p, r->vox, center, &extent);
    float lp[2][3];
    for (int j = 0; j < 2; ++j) {
        const float t = j == 0 ? r->t_near : r->t_far;
        for (int a = 0; a < 3; ++a) {
            const float pos = r->o[a] + r->d[a] * t;
            /* index_to_local_voxel_pos with clamping */
            const float l = (pos - (float)r->vox[a]) - 0.5f;
            lp[j][a] = fminf(fmaxf(l, -0.5f), 0.5f);
        }
    }
    /* get_deterministic_local_pi_po: face = largest |coordinate| (first of
     * equals), 2 axis for the + side, 2 axis + 1 for the - side; the point on
     * the lower-numbered face goes first */
    int face[2];
    for (int j = 0; j < 2; ++j) {
        float best = -2e20f;
        face[j] = 0;
        for (int a = 0; a < 3; ++a) {
            const float dist = fabsf(lp[j][a]);
            if (dist > best) {
                face[j] = lp[j][a] > 0.0f ? 2 * a : 2 * a + 1;
                best = dist;
            }
        }
    }
    const int swap = face[0] > face[1];
    const float dil = fminf(p->inference_min_extent_dilation, p->voxel_extent_dilation_outward);
    for (int j = 0; j < 2; ++j) {
        const float* src = lp[swap ? 1 - j : j];
        for (int a = 0; a < 3; ++a) {
            const float l = src[a] * (1.0f + dil);
            const float w = center[a] + l * extent;
            in[3 * j + a] = (w - sc->bmin[a]) / (sc->bmax[a] - sc->bmin[a]);
        }
    }
}

static void throughput_inputs(const nlo_tracer_scene* sc, const nlo_tracer_params* p, nlo_ray* r, float* in) {
    const float u0 = nlo_lcg_next(&r->rng);
    const float u1 = nlo_lcg_next(&r->rng);
    float wi[3];
    if (p->use_hg) sample_hg(u0, u1, r->d, p->hg_g, wi);
    else uniform_sphere(u0, u1, wi);
    float center[3], extent;
    voxel_of(sc, p, r->vox, center, &extent);
    for (int a = 0; a < 3; ++a) {
        in[a] = (center[a] - sc->bmin[a]) / (sc->bmax[a] - sc->bmin[a]);
        in[3 + a] = (-r->d[a] + 1.0f) * 0.5f;
        in[6 + a] = (wi[a] + 1.0f) * 0.5f;
        r->wi[a] = wi[a];
    }
    if (p->n_thr_inputs > 9) in[9] = (float)p->lod;
    r->pdf_inv = p->use_hg ? 1.0f / hg_pdf(dot3(r->d, wi), p->hg_g) : (float)(4.0f * 3.14159265358979323846);
}

/* check_visibility (non-stochastic unless configured); vis RR per :75-87 */
static void check(const nlo_tracer_params* p, nlo_ray* r, float v, uint32_t k, float max_rr_prob) {
    r->needed += 1;
    const int occluded = p->stochastic_threshold ? v >= nlo_lcg_next(&r->rng) : v >= r->thr;
    if (occluded) {
        r->state = NLO_HIT;
    } else if (p->apply_visibility_rr && k >= 5) {
        const float prob = fmaxf(v, max_rr_prob);
        const float u = nlo_lcg_next(&r->rng);
        if (u < prob) {
            for (int a = 0; a < 3; ++a) r->T[a] = r->T[a] / prob;
        } else {
            r->state = NLO_INACTIVE;
        }
    }
}

/* process_ray_bounce (:202-254) with get_throughput_inference (wavefront_neural_pt.cuh:256-273) */
static void bounce(const nlo_tracer_params* p, nlo_ray* r, const float* out, uint32_t b) {
    float th[3];
    for (int a = 0; a < 3; ++a) {
        float v = out[a];
        if (p->log_learning) v = (float)nlo_exp((double)v) - 1.0f;
        th[a] = fmaxf(0.0f, v);
    }
    for (int a = 0; a < 3; ++a) r->T[a] = r->T[a] * (th[a] * r->pdf_inv);
    int cont = r->T[0] != 0.0f || r->T[1] != 0.0f || r->T[2] != 0.0f;
    const float t_eps = (float)((double)r->t_near + 1e-4 * (double)(r->t_far - r->t_near));
    for (int a = 0; a < 3; ++a) {
        r->o[a] = r->o[a] + r->d[a] * t_eps;
        r->d[a] = r->wi[a];
    }
    r->t_near = r->t_far = -1.0f;
    if (p->apply_rr && cont && b >= p->rr_start_bounce) {
        /* rr_terminate (wavefront_neural_pt.cuh:57-73) */
        const float hi = fmaxf(r->T[0], fmaxf(r->T[1], r->T[2]));
        const float prob = (float)fmin(1.0, (double)hi);
        const float u = nlo_lcg_next(&r->rng);
        if (u < prob) {
            for (int a = 0; a < 3; ++a) r->T[a] = r->T[a] / prob;
        } else {
            cont = 0;
        }
    }
    r->state = cont ? NLO_ACTIVE : NLO_INACTIVE;
}

static void emit(const nlo_ray* r, float* illum, uint32_t* inferences, uint32_t* bounces) {
    for (int a = 0; a < 3; ++a) illum[(uint64_t)r->pixel * 3 + a] = r->illum[a];
    if (inferences) inferences[r->pixel] = r->needed;
    if (bounces) bounces[r->pixel] = r->bounce;
}

/* voxel_tracing: one sample per pixel.  illum float[n][3], inferences and
 * bounces (the bounce at which each path ended) uint32[n], nullable. */
void nlo_trace(const nlo_tracer_scene* sc, const nlo_camera* cam, const nlo_tracer_params* p, uint32_t sample_idx,
               uint32_t sample_offset, nlo_net_fn net, void* user, float* illum, uint32_t* inferences,
               uint32_t* bounces) {
    const uint32_t n = cam->width * cam->height;
    /* max_visibility_rr_prob_const (wavefront_neural_pt.cuh:347) */
    const float max_rr_prob = powf(1e-6f, 1.f / (float)(3 * p->grid_res));
    const uint32_t max_vis = (p->grid_res >> p->lod) * 3;
    const uint32_t nti = p->n_thr_inputs;
    nlo_ray* rays = (nlo_ray*)malloc(sizeof(nlo_ray) * (n ? n : 1));
    uint32_t* active = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* next = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* hit = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    float* vin = (float*)malloc(sizeof(float) * 6 * (n ? n : 1));
    float* vout = (float*)malloc(sizeof(float) * (n ? n : 1));
    float* tin = (float*)malloc(sizeof(float) * nti * (n ? n : 1));
    float* tout = (float*)malloc(sizeof(float) * 3 * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) {
        raygen(sc, cam, p, sample_idx, sample_offset, i, &rays[i]);
        active[i] = i;
    }
    uint32_t n_active = n;
    for (uint32_t b = 0; b < p->max_depth; ++b) {
        uint32_t n_hit = 0;
        const int last_bounce = b == p->max_depth - 1;
        for (uint32_t k = 0; k < max_vis; ++k) {
            for (uint32_t j = 0; j < n_active; ++j) {
                nlo_ray* r = &rays[active[j]];
                if (r->state == NLO_ACTIVE && !march(sc, r)) r->state = NLO_ESCAPED;
            }
            /* compact_for_visibility_and_accumulate */
            const int last_inference = k == max_vis - 1;
            uint32_t n_next = 0;
            for (uint32_t j = 0; j < n_active; ++j) {
                nlo_ray* r = &rays[active[j]];
                if (r->state == NLO_ACTIVE && !last_inference) {
                    next[n_next++] = active[j];
                } else if (r->state == NLO_HIT && !last_bounce) {
                    hit[n_hit++] = active[j];
                } else {
                    finalize(p, r, b);
                    emit(r, illum, inferences, bounces);
                }
            }
            memcpy(active, next, sizeof(uint32_t) * n_next);
            n_active = n_next;
            if (n_active == 0) break;
            for (uint32_t j = 0; j < n_active; ++j) visibility_inputs(sc, p, &rays[active[j]], vin + 6 * j);
            net(user, 0, n_active, vin, vout);
            for (uint32_t j = 0; j < n_active; ++j) check(p, &rays[active[j]], vout[j], k, max_rr_prob);
        }
        if (n_hit == 0) break;
        for (uint32_t j = 0; j < n_hit; ++j) throughput_inputs(sc, p, &rays[hit[j]], tin + (uint64_t)nti * j);
        net(user, 1, n_hit, tin, tout);
        n_active = 0;
        for (uint32_t j = 0; j < n_hit; ++j) {
            nlo_ray* r = &rays[hit[j]];
            bounce(p, r, tout + 3 * j, b);
            /* compact_for_next_bounce_and_accumulate */
            if (r->state == NLO_ACTIVE) {
                active[n_active++] = hit[j];
            } else {
                finalize(p, r, b);
                emit(r, illum, inferences, bounces);
            }
        }
    }
    free(rays); free(active); free(next); free(hit); free(vin); free(vout); free(tin); free(tout);
}

/* ---------------------------------------------------------- test hooks */
/* The throughput query's direction sampler (uniform sphere or HG around the
 * ray direction w_o) and its inverse pdf, as throughput_inputs uses them. */
void nlo_tracer_sample_dir(float u0, float u1, int hg, const float* w_o, float g, float* out, float* pdf_inv) {
    if (hg) {
        sample_hg(u0, u1, w_o, g, out);
        *pdf_inv = 1.0f / hg_pdf(dot3(w_o, out), g);
    } else {
        uniform_sphere(u0, u1, out);
        *pdf_inv = (float)(4.0f * 3.14159265358979323846);
    }
}

/* One raymarch_impl step on a ray state (state ACTIVE expected). */
int nlo_tracer_march(const nlo_tracer_scene* sc, nlo_ray* r) { return march(sc, r); }
