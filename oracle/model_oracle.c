/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg call this code, and only as the checker (or, for
 * the FP32 inference loop, as the reported CPU baseline).  The product never
 * links or loads it.
 *
 * CPU restatement of the tiny-cuda-nn (v1.7, as vendored in neural_lod) operators
 * on the neural-LoD path.  Index arithmetic that decides WHICH table entries are
 * read is done exactly as the reference does it (float32 position scaling,
 * uint32 hashing); the weighted sums, the MLP and the gradients are accumulated
 * in double precision.  Paths below are relative to rptr/ext/tiny-cuda-nn/.
 *
 *   grid level table      <- include/tiny-cuda-nn/encodings/grid.h:958-1022
 *                            (offsets), :195-204 (grid_scale, grid_resolution)
 *   grid_index / hash     <- grid.h:110-135, :170-187
 *   pos_fract             <- include/tiny-cuda-nn/common_device.h:438-450
 *   grid forward          <- grid.h:220-392 (kernel_grid, n-linear branch)
 *   grid backward         <- grid.h:394-517 (kernel_grid_backward)
 *   spherical harmonics   <- encodings/spherical_harmonics.h:46-151, written
 *                            from the closed forms of the real SH basis
 *   identity              <- encodings/identity.h:45-68
 *   composite padding     <- encodings/composite.h:189-199,381-386 (layout is
 *                            passed in by the caller as segments)
 *   fully fused MLP       <- src/fully_fused_mlp.cu:499-557 (forward, no biases),
 *                            :733-833 (backward), activations common_device.h:67-244
 *   losses                <- losses/binary_cross_entropy.h:7-43 (neural_lod),
 *                            losses/relative_l2.h:39-76
 *   Adam                  <- optimizers/adam.h:48-119 (float32, op by op)
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { SEG_GRID = 0, SEG_SH = 1, SEG_IDENTITY = 2 };
enum { GRID_HASH = 0, GRID_DENSE = 1, GRID_TILED = 2 };
enum { INTERP_LINEAR = 0, INTERP_NEAREST = 1, INTERP_SMOOTHSTEP = 2 };
enum { ACT_NONE = 0, ACT_RELU = 1, ACT_SIGMOID = 2, ACT_EXP = 3 };
enum { LOSS_BCE = 0, LOSS_REL_L2 = 1, LOSS_L2 = 2 };

#define MAX_LEVELS 32

/* One nested encoding of a Composite: which input columns it reads, which
 * encoded rows it writes, and where its parameters start. */
typedef struct {
    int32_t type;
    int32_t in_offset, n_in;
    int32_t out_offset, n_out, n_pad;
    /* grid */
    int32_t n_levels, n_features, log2_hashmap_size, base_resolution;
    float per_level_scale;
    int32_t grid_type, interp;
    int64_t param_offset;
    /* spherical harmonics */
    int32_t degree;
    /* identity */
    float scale, offset;
} nlo_seg;

typedef struct {
    int32_t width, n_hidden_layers, n_input, n_output, n_output_padded;
    int32_t activation, output_activation;
} nlo_net;

/* ------------------------------------------------------------ grid layout */

/* grid.h:195-204 with the host's exp2f/log2f (float) */
static float grid_scale(uint32_t level, float log2_scale, uint32_t base) {
    return exp2f((float)level * log2_scale) * (float)base - 1.0f;
}
static uint32_t grid_resolution(float scale) { return (uint32_t)ceilf(scale) + 1u; }

/* Per-level entry offsets (n_levels + 1), scales and resolutions. */
int64_t nlo_grid_levels(const nlo_seg* s, uint32_t* offsets, float* scales, uint32_t* res) {
    const float l2s = log2f(s->per_level_scale);
    uint32_t off = 0;
    for (int l = 0; l < s->n_levels; ++l) {
        float sc = grid_scale((uint32_t)l, l2s, (uint32_t)s->base_resolution);
        uint32_t r = grid_resolution(sc);
        const uint32_t max_params = 0xffffffffu / 2u;
        uint32_t params;
        if (powf((float)r, 3.0f) > (float)max_params) params = max_params;
        else params = r * r * r;
        params = (params + 7u) / 8u * 8u;
        if (s->grid_type == GRID_TILED) {
            uint32_t b = (uint32_t)s->base_resolution;
            if (b * b * b < params) params = b * b * b;
        } else if (s->grid_type == GRID_HASH) {
            uint32_t t = 1u << s->log2_hashmap_size;
            if (t < params) params = t;
        }
        if (offsets) offsets[l] = off;
        if (scales) scales[l] = sc;
        if (res) res[l] = r;
        off += params;
    }
    if (offsets) offsets[s->n_levels] = off;
    return (int64_t)off * s->n_features;
}

static uint32_t grid_index(int grid_type, uint32_t hashmap_size, uint32_t resolution, const uint32_t pg[3]) {
    uint32_t stride = 1, index = 0;
    for (int d = 0; d < 3 && stride <= hashmap_size; ++d) {
        index += pg[d] * stride;
        stride *= resolution;
    }
    if (grid_type == GRID_HASH && hashmap_size < stride) {
        /* coherent prime hash, primes {1, 2654435761, 805459861} */
        index = (pg[0] * 1u) ^ (pg[1] * 2654435761u) ^ (pg[2] * 805459861u);
    }
    return index % hashmap_size;
}

/* pos_fract: pos = x*scale + 0.5 as one fused multiply-add (the reference's
 * CUDA build contracts it), integer cell and fraction in float. */
static void pos_fract(float x, float scale, int interp, float* frac, uint32_t* cell) {
    float p = fmaf(x, scale, 0.5f);
    float fl = floorf(p);
    *cell = (uint32_t)(int)fl;
    float f = p - fl;
    if (interp == INTERP_SMOOTHSTEP) f = f * f * (3.0f - 2.0f * f);
    *frac = f;
}

/* ----------------------------------------------------- encoding forward */

#define SH_MAX 64

/* Real spherical harmonics (Condon-Shortley phase, tcnn ordering l*l+l+m) of
 * the unit vector (x, y, z), degrees 0..deg-1, from the associated Legendre
 * recurrence in double precision. */
static void sh_eval(int deg, double x, double y, double z, double* out) {
    /* P[l][m] = P_l^m(z) / (1-z^2)^{m/2} ("reduced" so that multiplying by
     * Re/Im (x+iy)^m gives the polynomial form) */
    double P[9][9];
    memset(P, 0, sizeof(P));
    P[0][0] = 1.0;
    for (int m = 1; m < deg; ++m) P[m][m] = -(2.0 * m - 1.0) * P[m - 1][m - 1];
    for (int m = 0; m + 1 < deg; ++m) P[m + 1][m] = (2.0 * m + 1.0) * z * P[m][m];
    for (int m = 0; m < deg; ++m)
        for (int l = m + 2; l < deg; ++l)
            P[l][m] = ((2.0 * l - 1.0) * z * P[l - 1][m] - (l + m - 1.0) * P[l - 2][m]) / (double)(l - m);
    double cr[9], ci[9]; /* Re, Im of (x + i y)^m */
    cr[0] = 1.0; ci[0] = 0.0;
    for (int m = 1; m < deg; ++m) {
        cr[m] = cr[m - 1] * x - ci[m - 1] * y;
        ci[m] = cr[m - 1] * y + ci[m - 1] * x;
    }
    for (int l = 0; l < deg; ++l) {
        for (int m = -l; m <= l; ++m) {
            int am = m < 0 ? -m : m;
            double fact = 1.0; /* (l-|m|)! / (l+|m|)! */
            for (int k = l - am + 1; k <= l + am; ++k) fact /= (double)k;
            double K = sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * fact);
            double v;
            if (m == 0) v = K * P[l][0];
            else if (m > 0) v = sqrt(2.0) * K * P[l][am] * cr[am];
            else v = sqrt(2.0) * K * P[l][am] * ci[am];
            out[l * l + l + m] = v;
        }
    }
}

/* Encoded features of one query: enc[0..enc_width) (double). */
static void encode_one(const nlo_seg* segs, int n_segs, const float* params, const float* in,
                       double* enc, int64_t* corner_idx /* optional, for backward */) {
    for (int si = 0; si < n_segs; ++si) {
        const nlo_seg* s = &segs[si];
        const float* x = in + s->in_offset;
        double* o = enc + s->out_offset;
        if (s->type == SEG_GRID) {
            uint32_t offs[MAX_LEVELS + 1];
            float scales[MAX_LEVELS];
            uint32_t res[MAX_LEVELS];
            nlo_grid_levels(s, offs, scales, res);
            const int F = s->n_features;
            for (int l = 0; l < s->n_levels; ++l) {
                const uint32_t hs = offs[l + 1] - offs[l];
                float fr[3];
                uint32_t cell[3];
                for (int d = 0; d < 3; ++d) pos_fract(x[d], scales[l], s->interp, &fr[d], &cell[d]);
                const float* table = params + s->param_offset + (int64_t)offs[l] * F;
                for (int f = 0; f < F; ++f) o[l * F + f] = 0.0;
                if (s->interp == INTERP_NEAREST) {
                    uint32_t idx = grid_index(s->grid_type, hs, res[l], cell);
                    for (int f = 0; f < F; ++f) o[l * F + f] = table[(int64_t)idx * F + f];
                    continue;
                }
                for (int c = 0; c < 8; ++c) {
                    double w = 1.0;
                    uint32_t pg[3];
                    for (int d = 0; d < 3; ++d) {
                        if ((c & (1 << d)) == 0) { w *= 1.0 - (double)fr[d]; pg[d] = cell[d]; }
                        else { w *= (double)fr[d]; pg[d] = cell[d] + 1u; }
                    }
                    uint32_t idx = grid_index(s->grid_type, hs, res[l], pg);
                    if (corner_idx) corner_idx[l * 8 + c] = (int64_t)idx;
                    for (int f = 0; f < F; ++f) o[l * F + f] += w * (double)table[(int64_t)idx * F + f];
                }
            }
            for (int p = 0; p < s->n_pad; ++p) o[s->n_out + p] = 0.0; /* grid.h:1041-1051 */
        } else if (s->type == SEG_SH) {
            /* pad rows of 1.0 come first (spherical_harmonics.h:58-63) */
            for (int p = 0; p < s->n_pad; ++p) o[p] = 1.0;
            float fx = x[0] * 2.f - 1.f, fy = x[1] * 2.f - 1.f, fz = x[2] * 2.f - 1.f;
            double sh[SH_MAX];
            sh_eval(s->degree, fx, fy, fz, sh);
            for (int k = 0; k < s->degree * s->degree; ++k) o[s->n_pad + k] = sh[k];
        } else { /* identity: x*scale + offset, pad rows 1.0 after (identity.h:63-66) */
            for (int k = 0; k < s->n_out; ++k) o[k] = (double)(x[k] * s->scale + s->offset);
            for (int p = 0; p < s->n_pad; ++p) o[s->n_out + p] = 1.0;
        }
    }
}

void nlo_encode(const nlo_seg* segs, int n_segs, int enc_width, const float* params, const float* in,
                int n_in, int64_t N, double* enc) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) encode_one(segs, n_segs, params, in + i * n_in, enc + i * enc_width, NULL);
}

/* --------------------------------------------------------------- MLP */

static double act_fwd(int a, double v) {
    switch (a) {
        case ACT_RELU: return v > 0.0 ? v : 0.0;
        case ACT_SIGMOID: return 1.0 / (1.0 + exp(-v));
        case ACT_EXP: return exp(v);
        default: return v;
    }
}
/* derivative given the post-activation value y (common_device.h:193-244) */
static double act_bwd(int a, double y) {
    switch (a) {
        case ACT_RELU: return y > 0.0 ? 1.0 : 0.0;
        case ACT_SIGMOID: return y * (1.0 - y);
        case ACT_EXP: return y;
        default: return 1.0;
    }
}

int64_t nlo_net_n_params(const nlo_net* n) {
    return (int64_t)n->width * n->n_input + (int64_t)(n->n_hidden_layers - 1) * n->width * n->width +
           (int64_t)n->n_output_padded * n->width;
}

/* Forward of one query: hidden[n_hidden][W] post-activation, out[n_output_padded]. */
static void mlp_one(const nlo_net* n, const float* w, const double* x, double* hidden, double* out) {
    const int W = n->width;
    const float* Wm = w;
    const double* cur = x;
    int fan_in = n->n_input;
    for (int layer = 0; layer < n->n_hidden_layers; ++layer) {
        double* h = hidden + (int64_t)layer * W;
        for (int r = 0; r < W; ++r) {
            double acc = 0.0;
            for (int k = 0; k < fan_in; ++k) acc += (double)Wm[(int64_t)r * fan_in + k] * cur[k];
            h[r] = act_fwd(n->activation, acc);
        }
        Wm += (int64_t)W * fan_in;
        fan_in = W;
        cur = h;
    }
    for (int r = 0; r < n->n_output_padded; ++r) {
        double acc = 0.0;
        for (int k = 0; k < W; ++k) acc += (double)Wm[(int64_t)r * W + k] * cur[k];
        out[r] = act_fwd(n->output_activation, acc);
    }
}

/* out: [N][n_output_padded]; hidden (optional): [N][n_hidden][W] */
void nlo_mlp_forward(const nlo_net* n, const float* w, const double* enc, int64_t N, double* out, double* hidden) {
    const int64_t hs = (int64_t)n->n_hidden_layers * n->width;
#pragma omp parallel
    {
        double* tmp = hidden ? NULL : (double*)malloc(sizeof(double) * hs);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < N; ++i)
            mlp_one(n, w, enc + i * n->n_input, hidden ? hidden + i * hs : tmp, out + i * n->n_output_padded);
        free(tmp);
    }
}

/* Backward of one query.  d_out: dL/d(output) [n_output_padded] (post
 * activation).  Accumulates dL/dW into gw (double, network param layout) and
 * writes dL/d(enc) into d_in [n_input]. */
static void mlp_bwd_one(const nlo_net* n, const float* w, const double* x, const double* hidden, const double* out,
                        const double* d_out, double* gw, double* d_in) {
    const int W = n->width, L = n->n_hidden_layers;
    double delta[256], delta_prev[256];
    /* output layer */
    int64_t off_out = (int64_t)W * n->n_input + (int64_t)(L - 1) * W * W;
    const double* h_last = hidden + (int64_t)(L - 1) * W;
    double dz_out[64];
    for (int r = 0; r < n->n_output_padded; ++r) dz_out[r] = d_out[r] * act_bwd(n->output_activation, out[r]);
    for (int r = 0; r < n->n_output_padded; ++r)
        for (int k = 0; k < W; ++k) gw[off_out + (int64_t)r * W + k] += dz_out[r] * h_last[k];
    for (int k = 0; k < W; ++k) {
        double acc = 0.0;
        for (int r = 0; r < n->n_output_padded; ++r) acc += (double)w[off_out + (int64_t)r * W + k] * dz_out[r];
        delta[k] = acc * act_bwd(n->activation, h_last[k]);
    }
    /* hidden matmuls, last to first */
    for (int layer = L - 1; layer >= 1; --layer) {
        int64_t off = (int64_t)W * n->n_input + (int64_t)(layer - 1) * W * W;
        const double* h_in = hidden + (int64_t)(layer - 1) * W;
        for (int r = 0; r < W; ++r)
            for (int k = 0; k < W; ++k) gw[off + (int64_t)r * W + k] += delta[r] * h_in[k];
        for (int k = 0; k < W; ++k) {
            double acc = 0.0;
            for (int r = 0; r < W; ++r) acc += (double)w[off + (int64_t)r * W + k] * delta[r];
            delta_prev[k] = acc * act_bwd(n->activation, h_in[k]);
        }
        memcpy(delta, delta_prev, sizeof(double) * W);
    }
    /* input layer */
    for (int r = 0; r < W; ++r)
        for (int k = 0; k < n->n_input; ++k) gw[(int64_t)r * n->n_input + k] += delta[r] * x[k];
    for (int k = 0; k < n->n_input; ++k) {
        double acc = 0.0;
        for (int r = 0; r < W; ++r) acc += (double)w[(int64_t)r * n->n_input + k] * delta[r];
        d_in[k] = acc;
    }
}

/* Serial over queries (deterministic double accumulation). */
void nlo_mlp_backward(const nlo_net* n, const float* w, const double* enc, const double* hidden, const double* out,
                      const double* d_out, int64_t N, double* gw, double* d_enc) {
    const int64_t hs = (int64_t)n->n_hidden_layers * n->width;
    memset(gw, 0, sizeof(double) * nlo_net_n_params(n));
    for (int64_t i = 0; i < N; ++i)
        mlp_bwd_one(n, w, enc + i * n->n_input, hidden + i * hs, out + i * n->n_output_padded,
                    d_out + i * n->n_output_padded, gw, d_enc + i * n->n_input);
}

/* Grid parameter gradients: for every grid segment, scatter dL/d(enc) with the
 * interpolation weights (grid.h:394-517).  genc: double [n_params_total] indexed
 * like the model parameters (only grid ranges are written; accumulated). */
void nlo_encode_backward(const nlo_seg* segs, int n_segs, int enc_width, const float* in, int n_in, int64_t N,
                         const double* d_enc, double* gparams) {
    for (int64_t i = 0; i < N; ++i) {
        const float* q = in + i * n_in;
        const double* g = d_enc + i * enc_width;
        for (int si = 0; si < n_segs; ++si) {
            const nlo_seg* s = &segs[si];
            if (s->type != SEG_GRID) continue;
            uint32_t offs[MAX_LEVELS + 1];
            float scales[MAX_LEVELS];
            uint32_t res[MAX_LEVELS];
            nlo_grid_levels(s, offs, scales, res);
            const int F = s->n_features;
            const float* x = q + s->in_offset;
            for (int l = 0; l < s->n_levels; ++l) {
                const uint32_t hs = offs[l + 1] - offs[l];
                float fr[3];
                uint32_t cell[3];
                for (int d = 0; d < 3; ++d) pos_fract(x[d], scales[l], s->interp, &fr[d], &cell[d]);
                double* tab = gparams + s->param_offset + (int64_t)offs[l] * F;
                const double* gl = g + s->out_offset + l * F;
                if (s->interp == INTERP_NEAREST) {
                    uint32_t idx = grid_index(s->grid_type, hs, res[l], cell);
                    for (int f = 0; f < F; ++f) tab[(int64_t)idx * F + f] += gl[f];
                    continue;
                }
                for (int c = 0; c < 8; ++c) {
                    double w = 1.0;
                    uint32_t pg[3];
                    for (int d = 0; d < 3; ++d) {
                        if ((c & (1 << d)) == 0) { w *= 1.0 - (double)fr[d]; pg[d] = cell[d]; }
                        else { w *= (double)fr[d]; pg[d] = cell[d] + 1u; }
                    }
                    uint32_t idx = grid_index(s->grid_type, hs, res[l], pg);
                    for (int f = 0; f < F; ++f) tab[(int64_t)idx * F + f] += w * gl[f];
                }
            }
        }
    }
}

/* Table entries (level-relative indices) touched by query 0 of `in` for a grid
 * segment: idx[level*8 + corner] — used by tests of the hashing. */
void nlo_grid_corners(const nlo_seg* s, const float* x, int64_t* idx) {
    double enc[MAX_LEVELS * 8];
    nlo_seg t = *s;
    t.in_offset = 0;
    t.out_offset = 0;
    /* the table values are irrelevant for the indices: use a zero table */
    int64_t np = nlo_grid_levels(s, NULL, NULL, NULL);
    float* zeros = (float*)calloc((size_t)np, sizeof(float));
    t.param_offset = 0;
    encode_one(&t, 1, zeros, x, enc, idx);
    free(zeros);
}

/* ---------------------------------------------------------- FP32 loop */

/* Straightforward FP32 encode + MLP inference (OpenMP over queries): the CPU
 * baseline loop.  out: float [N][n_output] (unpadded). */
void nlo_infer_f32(const nlo_seg* segs, int n_segs, const nlo_net* n, const float* params, const float* in, int n_in,
                   int64_t N, float* out) {
    const int W = n->width;
    const int64_t net_params = nlo_net_n_params(n);
    const float* enc_params = params + net_params;
#pragma omp parallel
    {
        float enc[1024], h0[256], h1[256];
        uint32_t offs[8][MAX_LEVELS + 1];
        float scales[8][MAX_LEVELS];
        uint32_t res[8][MAX_LEVELS];
        for (int si = 0; si < n_segs && si < 8; ++si)
            if (segs[si].type == SEG_GRID) nlo_grid_levels(&segs[si], offs[si], scales[si], res[si]);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            const float* q = in + i * n_in;
            for (int si = 0; si < n_segs; ++si) {
                const nlo_seg* s = &segs[si];
                const float* x = q + s->in_offset;
                float* o = enc + s->out_offset;
                if (s->type == SEG_GRID) {
                    const int F = s->n_features;
                    for (int l = 0; l < s->n_levels; ++l) {
                        const uint32_t hs = offs[si][l + 1] - offs[si][l];
                        float fr[3];
                        uint32_t cell[3];
                        for (int d = 0; d < 3; ++d) pos_fract(x[d], scales[si][l], s->interp, &fr[d], &cell[d]);
                        const float* table = enc_params + (s->param_offset - net_params) + (int64_t)offs[si][l] * F;
                        float acc[8] = {0};
                        for (int c = 0; c < 8; ++c) {
                            float w = 1.0f;
                            uint32_t pg[3];
                            for (int d = 0; d < 3; ++d) {
                                if ((c & (1 << d)) == 0) { w *= 1.0f - fr[d]; pg[d] = cell[d]; }
                                else { w *= fr[d]; pg[d] = cell[d] + 1u; }
                            }
                            uint32_t idx = grid_index(s->grid_type, hs, res[si][l], pg);
                            for (int f = 0; f < F; ++f) acc[f] += w * table[(int64_t)idx * F + f];
                        }
                        for (int f = 0; f < F; ++f) o[l * F + f] = acc[f];
                    }
                    for (int p = 0; p < s->n_pad; ++p) o[s->n_out + p] = 0.0f;
                } else if (s->type == SEG_SH) {
                    for (int p = 0; p < s->n_pad; ++p) o[p] = 1.0f;
                    double sh[SH_MAX];
                    sh_eval(s->degree, x[0] * 2.f - 1.f, x[1] * 2.f - 1.f, x[2] * 2.f - 1.f, sh);
                    for (int k = 0; k < s->degree * s->degree; ++k) o[s->n_pad + k] = (float)sh[k];
                } else {
                    for (int k = 0; k < s->n_out; ++k) o[k] = x[k] * s->scale + s->offset;
                    for (int p = 0; p < s->n_pad; ++p) o[s->n_out + p] = 1.0f;
                }
            }
            const float* Wm = params;
            const float* cur = enc;
            int fan_in = n->n_input;
            float* hbuf[2] = {h0, h1};
            for (int layer = 0; layer < n->n_hidden_layers; ++layer) {
                float* h = hbuf[layer & 1];
                for (int r = 0; r < W; ++r) {
                    float a = 0.0f;
                    for (int k = 0; k < fan_in; ++k) a += Wm[(int64_t)r * fan_in + k] * cur[k];
                    h[r] = (n->activation == ACT_RELU) ? (a > 0.0f ? a : 0.0f) : a;
                }
                Wm += (int64_t)W * fan_in;
                fan_in = W;
                cur = h;
            }
            for (int r = 0; r < n->n_output; ++r) {
                float a = 0.0f;
                for (int k = 0; k < W; ++k) a += Wm[(int64_t)r * W + k] * cur[k];
                out[i * n->n_output + r] = (float)act_fwd(n->output_activation, a);
            }
        }
    }
}

/* ------------------------------------------------------------- losses */

/* pred: [N][stride] (padded outputs), target: [N][dims]; values/grads: [N][stride].
 * grads include loss_scale and 1/n_total like the reference. */
void nlo_loss(int type, const double* pred, const float* target, int64_t N, int stride, int dims, double loss_scale,
              double* values, double* grads) {
    const double n_total = (double)N * dims;
    for (int64_t i = 0; i < N; ++i) {
        for (int d = 0; d < stride; ++d) {
            int64_t k = i * stride + d;
            if (d >= dims) { values[k] = 0.0; grads[k] = 0.0; continue; }
            double t = target[i * dims + d];
            double p = pred[k];
            if (type == LOSS_BCE) {
                /* binary_cross_entropy.h:33-34 clamps the float prediction to [1e-5f, 1 - 1e-5f] */
                const float eps = 1e-5f;
                const float pf = fminf(fmaxf((float)p, eps), 1.0f - eps);
                p = (double)pf;
                values[k] = -(t * log(p) + (1.0 - t) * log(1.0 - p)) / n_total;
                grads[k] = loss_scale * (-(t - p) / (p - p * p)) / n_total;
            } else if (type == LOSS_REL_L2) {
                double den = p * p + 0.01;
                double diff = p - t;
                values[k] = diff * diff / den / n_total;
                grads[k] = loss_scale * (2.0 * diff / den) / n_total;
            } else {
                double diff = p - t;
                values[k] = diff * diff / n_total;
                grads[k] = loss_scale * (2.0 * diff) / n_total;
            }
        }
    }
}

/* -------------------------------------------------------------- Adam */

typedef struct nlo_adam {
    float lr, beta1, beta2, eps, l2_reg;
    float relative_decay, absolute_decay, clipping, non_matrix_lr_factor;
    int optimize_matrix, optimize_non_matrix, adabound;
} nlo_adam;

/* adam_step (optimizers/adam.h:48-119) in float32, one parameter at a time;
 * AdamOptimizer::step (:150-178) supplies the AdaBound bounds from the
 * optimizer's step count after its increment (opt_step). */
void nlo_adam_step(int64_t n, int64_t n_matrix, float loss_scale, const nlo_adam* a, uint32_t opt_step, float* w,
                   const float* grad, float* m, float* v, uint32_t* steps) {
    float lo = 0.0f, hi = FLT_MAX;
    if (a->adabound) {
        lo = 0.1f - 0.1f / ((1.0f - a->beta2) * (float)opt_step + 1.0f);
        hi = 0.1f + 0.1f / ((1.0f - a->beta2) * (float)opt_step);
    }
    for (int64_t i = 0; i < n; ++i) {
        float g = grad[i] / loss_scale;
        if (i >= n_matrix) {
            if (!a->optimize_non_matrix || g == 0.0f) continue;
        } else if (!a->optimize_matrix) {
            continue;
        }
        const float wf = w[i];
        if (i < n_matrix) g += a->l2_reg * wf;
        const float g2 = g * g;
        m[i] = a->beta1 * m[i] + (1.0f - a->beta1) * g;
        v[i] = a->beta2 * v[i] + (1.0f - a->beta2) * g2;
        float lr = a->lr;
        if (i >= n_matrix) lr *= a->non_matrix_lr_factor;
        const uint32_t t = ++steps[i];
        lr *= sqrtf(1.0f - powf(a->beta2, (float)t)) / (1.0f - powf(a->beta1, (float)t));
        const float eff = fminf(fmaxf(lr / (sqrtf(v[i]) + a->eps), lo), hi);
        const float decayed = (1.0f - a->relative_decay * lr) * wf - copysignf(a->absolute_decay * lr, wf);
        float nw = decayed - eff * m[i];
        if (a->clipping != 0.0f) nw = fminf(fmaxf(nw, -a->clipping), a->clipping);
        w[i] = nw;
    }
}

/* EmaOptimizer::step (optimizers/ema.h:45-62 and :95-130), full precision:
 * t = the nested optimizer's step count after its step. */
void nlo_ema_step(int64_t n, float decay, uint32_t t, const float* w, float* ema) {
    const float debias_old = 1.0f - (float)pow((double)decay, (double)t - 1.0);
    const float debias_new = 1.0f / (1.0f - (float)pow((double)decay, (double)t));
    for (int64_t i = 0; i < n; ++i) ema[i] = (ema[i] * decay * debias_old + w[i] * (1.0f - decay)) * debias_new;
}

int nlo_omp_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
