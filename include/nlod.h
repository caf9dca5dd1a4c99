/*
 * nlod.h — C ABI of the MI355X (gfx950) neural-LoD hot path.
 *
 * Every function returns an int status (NLOD_OK == 0); on failure the message
 * of the calling thread's last error is available from nlod_last_error().
 * No C++ exception crosses this boundary.  Device buffers passed in are owned
 * by the caller; model and trainer handles own their weights and workspaces.
 * Calls taking a hipStream_t (passed as void*) are asynchronous and ordered on
 * that stream unless documented otherwise.  Handles are not thread-safe.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the neural_lod repository).
 */
#ifndef NLOD_H
#define NLOD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */

enum nlod_status {
    NLOD_OK = 0,
    NLOD_ERR_INVALID_ARGUMENT = 1, /* bad pointer, size, shape or flag (tcnn CHECK_THROW) */
    NLOD_ERR_CONFIG = 2,           /* malformed or unsupported JSON configuration */
    NLOD_ERR_UNSUPPORTED = 3,      /* valid request this build does not implement */
    NLOD_ERR_HIP = 4,              /* a HIP runtime call failed */
    NLOD_ERR_OUT_OF_MEMORY = 5,
    NLOD_ERR_STATE = 6             /* call out of order (e.g. backward without forward) */
};

/* Message of the calling thread's most recent failure ("" if none).
 * Replaces the std::runtime_error texts of CHECK_THROW / checkCuda
 * (rptr/cuda/render_cuda.h:8-16, tiny-cuda-nn/common.h). */
const char* nlod_last_error(void);

/* Library version string. */
const char* nlod_version(void);

/* -------------------------------------------------------------- voxelizer */

/* Voxelization parameters.  Mirrors struct voxinfo
 * (neural_lod/cuda_voxelizer/src/util.h:57-76). */
typedef struct nlod_voxinfo {
    float bbox_min[3];
    float bbox_max[3];
    uint32_t gridsize[3];
    uint32_t reserved0;
    uint64_t n_triangles;
    float unit[3]; /* (bbox_max - bbox_min) / gridsize, per axis, in float */
    uint32_t reserved1;
} nlod_voxinfo;

enum nlod_vox_flags {
    NLOD_VOX_SURFACE = 0,      /* conservative surface voxelization (voxelize.cu:58-190) */
    NLOD_VOX_SOLID = 1u << 0,  /* parity solid fill (voxelize_solid.cu:73-146, -solid) */
    NLOD_VOX_MORTON = 1u << 1, /* Morton-ordered bits (-o morton); needs a power-of-two grid */
    NLOD_VOX_PAD_EPS = 1u << 2,/* pad the cube by 1e-4 of its extent (main.cpp:45,209-215, -pad) */
    NLOD_VOX_ACCUMULATE = 1u << 3, /* do not clear the table first */
    /* Kernel choice (same bits either way).  Default: LDS-tiled for row-major
     * tables of >= 4096 triangles with at most 16384 tiles, else direct. */
    NLOD_VOX_DIRECT = 1u << 4, /* one triangle per lane, global atomics */
    NLOD_VOX_TILED = 1u << 5   /* LDS-tiled whenever the table allows it, at any triangle count */
};

/* Number of uint32 words of a res^3 bit table: ceil(res^3 / 32)
 * (main.cpp:220). */
int nlod_vox_table_words(uint32_t res, uint64_t* words_out);

/* Cubic voxelization box from a mesh box, optionally padded by 1e-4 of its
 * extent.  Replaces createMeshBBCube (util.h:87-112) + the voxinfo
 * constructor (util.h:63-68) as called at main.cpp:214-216.  Host only. */
int nlod_voxinfo_from_bbox(const float bbox_min[3], const float bbox_max[3], uint32_t res,
                           uint32_t flags, uint64_t n_triangles, nlod_voxinfo* info_out);

/* Per-vertex bounding box of a device triangle soup float[n][9] (exact
 * min/max; replaces trimesh2 need_bbox, TriMesh_bounding.cc:18-31).
 * Synchronizes the stream to return the result to the host. */
int nlod_mesh_bbox(void* stream, const float* d_triangles, uint64_t n_triangles,
                   float bbox_min_out[3], float bbox_max_out[3]);

/* Voxelize a device triangle soup float[n][9] (x0 y0 z0 x1 ... z2 per
 * triangle) into the device table d_table of nlod_vox_table_words(G) words.
 * Clears the table first unless NLOD_VOX_ACCUMULATE.  Bit 31-(i%32) of word
 * i/32 holds voxel i = x + y*G + z*G*G.  Results are bit-identical to the
 * reference CPU voxelizer (cpu_voxelizer.cpp:39-193 / :252-328).
 * Replaces void voxelize(const voxinfo&, float*, unsigned*, bool, bool)
 * (voxelize.cu:192) and voxelize_solid(...) (voxelize_solid.cu:148).
 * Asynchronous. */
int nlod_voxelize(void* stream, const nlod_voxinfo* info, const float* d_triangles,
                  uint32_t* d_table, uint32_t flags);

/* voxelize(mesh, res): bounding box, cube, then nlod_voxelize.  Writes the
 * voxinfo used to info_out (may be NULL).  Synchronizes once (bounding box),
 * the voxelization itself is asynchronous.  Replaces the GPU branch of
 * main.cpp:199-258. */
int nlod_voxelize_mesh(void* stream, const float* d_triangles, uint64_t n_triangles, uint32_t res,
                       uint32_t flags, uint32_t* d_table, nlod_voxinfo* info_out);

/* nlod_voxelize that allocates nothing: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * flags and bits as nlod_voxelize; the LDS-tiled path takes its scratch from a
 * fixed device array of 32 MiB per GPU (held until the process ends) instead
 * of a library allocation.  A job whose tiled path needs more scratch than
 * that fails with NLOD_ERR_INVALID_ARGUMENT (about 700k triangles at 256^3,
 * 300k at 1024^3); NLOD_VOX_DIRECT needs no scratch at any size.  Every call
 * on one GPU shares its array: calls that overlap on one GPU (on different
 * streams or threads) race and must be serialized by the caller.
 * Asynchronous. */
int nlod_voxelize_noalloc(void* stream, const nlod_voxinfo* info, const float* d_triangles, uint32_t* d_table,
                          uint32_t flags);

/* ----------------------------------------- network with input encoding */

typedef struct nlod_model nlod_model;
typedef struct nlod_ctx nlod_ctx;

enum nlod_precision {
    NLOD_PRECISION_FP32 = 0, /* plain form: FP32 throughout */
    NLOD_PRECISION_FP16 = 1  /* FP16 MFMA GEMMs with FP32 accumulation; FP32 grid interpolation */
};

/* EGradientMode (tiny-cuda-nn include/tiny-cuda-nn/object.h:115-119) */
enum nlod_grad_mode { NLOD_GRAD_IGNORE = 0, NLOD_GRAD_OVERWRITE = 1, NLOD_GRAD_ACCUMULATE = 2 };

typedef struct nlod_model_info {
    uint64_t n_params;         /* network + encoding parameters (float) */
    uint64_t n_network_params; /* leading part of the parameter vector */
    uint32_t input_width;      /* floats per input row */
    uint32_t encoded_width;    /* padded encoding width = network input width */
    uint32_t padded_output_width;
    uint32_t output_width;
    uint32_t network_width;
    uint32_t n_hidden_layers;
    uint32_t precision;
    uint32_t reserved;
} nlod_model_info;

/* Model from tcnn JSON (encoding + network).  Replaces
 * tcnn::NetworkWithInputEncoding<T>(n_dims_to_encode, n_output_dims, encoding, network)
 * (tiny-cuda-nn include/tiny-cuda-nn/network_with_input_encoding.h:53-54),
 * as built by rptr/cuda/neural-lod/neural_lod_learning_init.cu:663-664,898-899.
 * Encodings: Composite, RepeatedComposite, HashGrid/Grid/DenseGrid/TiledGrid,
 * SphericalHarmonics, Identity.  Network: FullyFusedMLP (widths 16-128, no
 * biases), hidden ReLU/None, output None/ReLU/Sigmoid/Exponential. */
int nlod_model_create(uint32_t n_input_dims, uint32_t n_output_dims, const char* encoding_json,
                      const char* network_json, int precision, nlod_model** model_out);
int nlod_model_destroy(nlod_model* model);

/* n_params / widths (object.h:362-367, network_with_input_encoding.h:130-165). */
int nlod_model_get_info(const nlod_model* model, nlod_model_info* info_out);

/* JSON description of the flattened layout (segments, offsets, level tables).
 * Writes at most buf_len bytes (NUL-terminated); *needed gets the full size. */
int nlod_model_describe(const nlod_model* model, char* buf, size_t buf_len, size_t* needed);

/* tcnn Trainer initialisation (trainer.h:53-92: pcg32 from seed_seq{seed},
 * Xavier-uniform matrices, grid U[-1e-4,1e-4]) into a host float[n_params]. */
int nlod_model_init_params(const nlod_model* model, uint32_t seed, float* params_host);

/* Binds device parameter vectors (float[n_params], caller-owned, must stay
 * valid): training params and inference params (NULL = same).  Replaces
 * set_params(params, inference_params, gradients) (object.h:73-79).  For the
 * FP16 form this also repacks the FP16 MFMA weight copies on `stream`; call it
 * again after the parameters change. */
int nlod_model_set_params(nlod_model* model, void* stream, const float* d_params, const float* d_inference_params);

/* Inference: d_in float[n][input_width] -> d_out float[n][output_width].  Any
 * n (no multiple-of-128 restriction).  Replaces DifferentiableObject::inference
 * (object.h:147-169) as called from wavefront_neural_throughput_visibility_lod.cu:595,662. */
int nlod_model_inference(nlod_model* model, void* stream, uint32_t n, const float* d_in, float* d_out);

/* Forward with saved context (object.h:182-203); d_out may be NULL. */
int nlod_model_forward(nlod_model* model, void* stream, uint32_t n, const float* d_in, float* d_out,
                       int use_inference_params, int prepare_input_gradients, nlod_ctx** ctx_out);

/* Backward (object.h:223-259): d_dL_dout float[n][output_width] ->
 * d_dL_dparams float[n_params] (per grad_mode) and, if non-NULL,
 * d_dL_dinput float[n][input_width] (SH/identity encodings only). */
int nlod_model_backward(nlod_model* model, void* stream, nlod_ctx* ctx, const float* d_in, const float* d_dL_dout,
                        float* d_dL_dparams, float* d_dL_dinput, int grad_mode);
int nlod_ctx_destroy(nlod_ctx* ctx);

/* Encoding alone (tcnn Encoding::inference, encoding.h:59-88): encoded
 * features as SoA float[encoded_width][n] (row r of query i at r*n + i). */
int nlod_model_encode(nlod_model* model, void* stream, uint32_t n, const float* d_in, float* d_encoded,
                      int use_inference_params);
/* Grid parameter gradients from dL/d(encoded) SoA float[encoded_width][n]
 * (GridEncoding::backward_impl, grid.h:1102-1193). */
int nlod_model_encode_backward(nlod_model* model, void* stream, uint32_t n, const float* d_in,
                               const float* d_dL_dencoded, float* d_dL_dparams, int grad_mode);

/* FP16 models: the encoding pass alone, writing FP16 features in the layout
 * the MFMA kernels read, [ceil(n/16)][K/2][16][2] (16-query tiles of feature
 * pairs; K = encoded width padded to 32),
 * n_bytes from nlod_model_feature_bytes.  For profiling and for callers that
 * reuse one encoding across several networks; no counterpart in the reference. */
int nlod_model_feature_bytes(const nlod_model* model, uint32_t n, uint64_t* bytes_out);
int nlod_model_encode_fragments(nlod_model* model, void* stream, uint32_t n, const float* d_in, void* d_features);
/* FP16 models: the fused MLP alone on features that nlod_model_encode_fragments
 * wrote, -> d_out float[n][output_width] (inference parameters).  The second
 * half of nlod_model_inference, for profiling and for callers that evaluate
 * one encoding once; no counterpart in the reference. */
int nlod_model_mlp_fragments(nlod_model* model, void* stream, uint32_t n, const void* d_features, float* d_out);

/* ---------------------------------------------------------------- trainer */

typedef struct nlod_trainer nlod_trainer;

enum nlod_trainer_flags {
    NLOD_TRAINER_GRAPH = 1u << 0,  /* capture each batch size's step in a hipGraph (after one eager step) */
    NLOD_TRAINER_UNFUSED = 1u << 1 /* FP16: separate forward / loss / backward kernels instead of the fused pass */
};

/* Trainer<float, T, T>(model, create_optimizer(json), create_loss(json), seed)
 * (tiny-cuda-nn include/tiny-cuda-nn/trainer.h:53-60; src/optimizer.cu:47-80;
 * src/loss.cu:47-80), as built at rptr/cuda/neural-lod/neural_lod_learning_init.cu:655-668,890-903.
 * Optimizers: Adam (optimizers/adam.h), ExponentialDecay{nested}, Ema{nested}.
 * Losses: L2, RelativeL2, BinaryCrossEntropy.  Parameters are initialised as
 * nlod_model_init_params(seed) and bound to the model (the model must outlive
 * the trainer; its parameters are the trainer's while it lives).
 * loss_scale_json (NULL = defaults) replaces the reference's static scale of 128
 * (trainer.h:171) by dynamic loss scaling: {"initial_scale": 1024, "dynamic": true,
 * "growth_factor": 2, "backoff_factor": 0.5, "growth_interval": 2000,
 * "min_scale": 1, "max_scale": 2^32}.  A step whose FP64 gradient check finds a
 * non-finite value is skipped (no optimizer step) and the scale backs off. */
int nlod_trainer_create(nlod_model* model, const char* optimizer_json, const char* loss_json,
                        const char* loss_scale_json, uint32_t seed, uint32_t flags, nlod_trainer** trainer_out);
int nlod_trainer_destroy(nlod_trainer* trainer);

/* One training step (Trainer::training_step, trainer.h:161-186, with the
 * optimizer step): d_in float[n][input_width], d_target float[n][output_width].
 * Asynchronous on `stream`; no host synchronisation inside. */
int nlod_trainer_step(nlod_trainer* trainer, void* stream, uint32_t n, const float* d_in, const float* d_target);

typedef struct nlod_trainer_status {
    double loss;          /* loss of the last step (Trainer::loss, trainer.h:197-199), FP64 sum */
    double grad_norm;     /* L2 norm of the last step's unscaled gradient (FP64) */
    float loss_scale;     /* scale the next step will use */
    float learning_rate;  /* learning rate of the last optimizer step */
    uint32_t step;        /* optimizer steps taken */
    uint32_t skipped_steps;
    uint32_t last_step_skipped;
    uint32_t reserved;
} nlod_trainer_status;
/* State after the last step; synchronizes with the trainer's work. */
int nlod_trainer_get_status(nlod_trainer* trainer, nlod_trainer_status* status_out);

/* Device pointers owned by the trainer: FP32 master parameters
 * (params_full_precision, trainer.h:217), inference parameters (the EMA weights
 * when an Ema optimizer is configured, trainer.h:322-325) and gradients. */
int nlod_trainer_params(nlod_trainer* trainer, float** d_params, float** d_inference_params, float** d_gradients);
/* Copies one of the trainer's FP32 vectors (0 master parameters, 1 inference
 * parameters, 2 gradients, 3 first moments, 4 second moments) into the
 * caller's device buffer float[n_params], ordered after the trainer's work. */
int nlod_trainer_copy(nlod_trainer* trainer, void* stream, int which, float* d_dst);
/* set_params_full_precision (trainer.h:231-245): copies d_params into the
 * master and inference parameters. */
int nlod_trainer_set_params(nlod_trainer* trainer, void* stream, const float* d_params);

/* serialize / deserialize (trainer.h:267-307) as a binary snapshot: the
 * inference parameters, plus master parameters, moments, per-parameter step
 * counts and scaler state when with_optimizer_state.  buf NULL: size query. */
int nlod_trainer_serialize(nlod_trainer* trainer, int with_optimizer_state, void* buf, size_t buf_len,
                           size_t* needed);
int nlod_trainer_deserialize(nlod_trainer* trainer, const void* buf, size_t len);

/* The weight file of the reference (Trainer::serialize without optimizer
 * state, trainer.h:267-278, as write_weights_to_disk writes it,
 * neural_lod_learning_ui.cu:916-924): JSON text
 * {"n_params":N,"params_binary":{"bytes":[...],"subtype":null},"params_type":T}
 * holding the inference parameters in tcnn's order, in the network precision
 * as the reference's Trainer<float, T, T> stores them: T "float" (4 bytes per
 * parameter) for FP32 models, "__half" (2 bytes, rounded to nearest even) for
 * FP16 models.  buf NULL: *needed = size including the terminating NUL.
 * Synchronizes the trainer's stream. */
int nlod_trainer_weights_json(nlod_trainer* trainer, char* buf, size_t buf_len, size_t* needed);
/* Loads such a file (load_weights_from_disk, neural_lod_learning_ui.cu:
 * 935-947; Trainer::deserialize, trainer.h:280-307): params_type "float" or
 * "__half", bytes in the {"bytes": [...]} text form; an "optimizer" entry is
 * ignored.  Sets the master and inference parameters.  Synchronizes. */
int nlod_trainer_load_weights_json(nlod_trainer* trainer, void* stream, const char* json, size_t len);

/* Loss evaluation alone (Loss::evaluate with stride == dims):
 * values (nullable) and loss_scale * dL/dpred per element, n_total = n_elements.
 * loss_type: 0 L2 (losses/l2.h), 1 RelativeL2 (losses/relative_l2.h:39-76),
 * 2 BinaryCrossEntropy (losses/binary_cross_entropy.h:7-43). */
int nlod_loss_evaluate(void* stream, int loss_type, uint64_t n_elements, float loss_scale, const float* d_pred,
                       const float* d_target, float* d_values, float* d_gradients);

/* Adam hyperparameters (optimizers/adam.h:214-262). */
typedef struct nlod_adam_config {
    float learning_rate, beta1, beta2, epsilon, l2_reg;
    float relative_decay, absolute_decay, clipping_magnitude, non_matrix_learning_rate_factor;
    int optimize_matrix_params, optimize_non_matrix_params, adabound;
} nlod_adam_config;

/* One Adam step on caller buffers (adam_step, optimizers/adam.h:48-119;
 * optimizer_step = the optimizer's step count after this step, used by
 * AdaBound).  Gradients are divided by loss_scale; the first n_matrix
 * parameters are matrix weights.  Synchronizes the stream. */
int nlod_adam_step(void* stream, const nlod_adam_config* config, uint64_t n, uint64_t n_matrix, float loss_scale,
                   uint32_t optimizer_step, float* d_weights, float* d_gradients, float* d_first_moments,
                   float* d_second_moments, uint32_t* d_param_steps);

/* ------------------------------------------------------- wavefront queries */

typedef struct nlod_wavefront nlod_wavefront;

/* Engine for the per-wavefront queries of the neural LoD path tracer
 * (rptr/cuda/neural-lod/wavefront_neural_throughput_visibility_lod.cu:464-720):
 * per ray segment the visibility network, then, where the visibility reaches
 * `threshold` (check_visibility, :597-610), the throughput network (:656-662).
 * Both models must share a precision; their parameters must be set.  FP16
 * models of width 32/64 with one feature count run one persistent fused
 * launch per wavefront; otherwise the plain launches per network. */
int nlod_wavefront_create(nlod_model* visibility, nlod_model* throughput, float threshold, nlod_wavefront** out);
int nlod_wavefront_destroy(nlod_wavefront* wavefront);

/* Autotuner: times the launch shapes (split launches; fused kernel with 1..max
 * blocks per CU) once per wavefront-size bucket up to max_n queries and keeps
 * the fastest.  Synchronizes the stream.  Untuned engines use the fused kernel
 * at 1 block per CU where it applies. */
int nlod_wavefront_tune(nlod_wavefront* wavefront, void* stream, uint32_t max_n);
/* The shape the engine picks for n queries (0 = split launches, k = fused kernel, k blocks per CU). */
int nlod_wavefront_choice(const nlod_wavefront* wavefront, uint32_t n, int* shape_out);
int nlod_wavefront_max_shape(const nlod_wavefront* wavefront, int* max_shape_out);

/* One wavefront of n segments: d_vis_in float[n][vis inputs], d_thr_in
 * float[n][thr inputs] -> d_vis_out float[n][vis outputs], d_thr_out
 * float[n][thr outputs] (0 where the visibility is below the threshold).
 * Asynchronous.  _shape forces a launch shape (0 split, k > 0 fused). */
int nlod_wavefront_run(nlod_wavefront* wavefront, void* stream, uint32_t n, const float* d_vis_in,
                       const float* d_thr_in, float* d_vis_out, float* d_thr_out);
int nlod_wavefront_run_shape(nlod_wavefront* wavefront, void* stream, uint32_t n, const float* d_vis_in,
                             const float* d_thr_in, float* d_vis_out, float* d_thr_out, int shape);
/* As nlod_wavefront_run with one threshold per segment, d_thresholds float[n]
 * (neural_result.voxel_threshold, filled by the march, raymarch_impl,
 * rptr/cuda/wavefront_neural_pt.cuh:124-136, and compared as visibility >=
 * threshold in check_visibility, wavefront_neural_throughput_visibility_lod.cu:
 * 320-325; nlod_lod_grid_march_thresholds gives these values).  shape -1 =
 * the engine's choice, 0 = split launches, k > 0 = fused kernel. */
int nlod_wavefront_run_thresholds(nlod_wavefront* wavefront, void* stream, uint32_t n, const float* d_vis_in,
                                  const float* d_thr_in, const float* d_thresholds, float* d_vis_out,
                                  float* d_thr_out, int shape);

/* A sequence of n_batches wavefronts (counts on the host) stored back to back
 * in the buffers.  use_graph: the sequence is captured in a hipGraph on the
 * first call with these sizes and buffers (that call runs eagerly) and
 * replayed by later calls. */
int nlod_wavefront_run_batches(nlod_wavefront* wavefront, void* stream, uint32_t n_batches, const uint32_t* counts,
                               const float* d_vis_in, const float* d_thr_in, float* d_vis_out, float* d_thr_out,
                               int use_graph);
/* nlod_wavefront_run_batches with one threshold per segment, d_thresholds
 * float[sum of counts] stored back to back like the other buffers (see
 * nlod_wavefront_run_thresholds).  The graph is keyed by this buffer too. */
int nlod_wavefront_run_batches_thresholds(nlod_wavefront* wavefront, void* stream, uint32_t n_batches,
                                          const uint32_t* counts, const float* d_vis_in, const float* d_thr_in,
                                          const float* d_thresholds, float* d_vis_out, float* d_thr_out,
                                          int use_graph);

/* ------------------------------------------ training samples from a mesh */

/* Training batches for the two LoD networks, drawn on the GPU from a triangle
 * mesh.  They replace the Vulkan ray-query round trip that produces the
 * (inputs, targets) handed to training_step (neural_lod_learning.cu:588-598,
 * 686-694): fill_ray_queries_visibility / _throughput, the ray queries of
 * rptr/vulkan/raygen.rgen:505-730 and record_ray_query_results_* /
 * copy_valid_results_throughput_optim (neural_lod_learning_kernels.cuh:
 * 100-330, 364-585).
 *
 * The LoD pyramid: level l is the (res >> l)^3 grid over the cubic mesh box
 * (no padding); level 0 is occupied where the surface voxelization is, level
 * l where any of its 2^l-block of level-0 voxels is.  Every sample is a pure
 * function of (seed, index): voxel and level uniform, directions uniform on
 * the sphere (a simplification of the reference's dilated-box sampler). */
typedef struct nlod_labeler nlod_labeler;

/* Builds the pyramid and, per occupied voxel of each level, the triangles
 * whose clamped voxel range contains it.  res must be divisible by
 * 2^(n_lods - 1); n_lods in [1, 8].  d_triangles float[n][9] stays
 * caller-owned and must outlive the labeler.  Synchronizes the stream (the
 * set-up runs on the host). */
int nlod_labeler_create(void* stream, const float* d_triangles, uint64_t n_triangles, uint32_t res,
                        uint32_t n_lods, nlod_labeler** out);
int nlod_labeler_destroy(nlod_labeler* labeler);
/* Grid box, occupied voxels and list entries of one level (outputs nullable). */
int nlod_labeler_info(const nlod_labeler* labeler, uint32_t lod, nlod_voxinfo* grid_out,
                      uint32_t* n_occupied_out, uint64_t* n_list_entries_out);
/* Visibility samples first .. first + n - 1 at level 0 (raygen.rgen:690-730:
 * origin uniform in an occupied voxel, nearest hit, label 1 when it lies in
 * the voxel): d_inputs float[n][6] = [position in the box (0..1)^3,
 * (direction + 1) / 2], d_labels float[n].  Asynchronous. */
int nlod_labeler_visibility(const nlod_labeler* labeler, void* stream, uint64_t seed, uint64_t first, uint32_t n,
                            float* d_inputs, float* d_labels);
/* Throughput samples (raygen.rgen:505-688): level, voxel, outgoing wo and
 * light direction wi per sample; spp paths enter the voxel along -wo, add
 * next-event estimates along wi at up to max_depth hits inside the voxel and
 * continue by cosine-weighted bounces off two-sided Lambertian triangles
 * whose albedo is drawn from albedo_seed.  d_inputs float[n][10] = [voxel
 * centre in the box, (wo + 1) / 2, (wi + 1) / 2, level], d_targets float[n][3]
 * = mean RGB throughput of the valid paths (first ray hits inside the voxel),
 * d_valid float[n] = 1 when any path was valid.  Asynchronous. */
int nlod_labeler_throughput(const nlod_labeler* labeler, void* stream, uint64_t seed, uint64_t first, uint32_t n,
                            uint32_t spp, uint32_t max_depth, uint64_t albedo_seed, float* d_inputs,
                            float* d_targets, float* d_valid);

/* ------------------------------------------ LoD occupancy pyramid, ray marching */

/* A stack of occupancy bit tables in the voxelizer's layout: level 0 is a
 * copy of d_table (a G^3 voxelization described by info), level l + 1 the OR
 * of 2x2x2 blocks of level l.  Replaces the per-LoD nanovdb grids built from
 * the downsampled point cloud (neural_lod_learning_init.cu:201-295).
 * G must be divisible by 2^(n_lods - 1); n_lods in [1, 10].  Asynchronous. */
typedef struct nlod_lod_grid nlod_lod_grid;
int nlod_lod_grid_create(void* stream, const nlod_voxinfo* info, const uint32_t* d_table, uint32_t n_lods,
                         nlod_lod_grid** out);
int nlod_lod_grid_destroy(nlod_lod_grid* grid);
/* Device bit table and box of one level (outputs nullable; the table stays
 * owned by the grid). */
int nlod_lod_grid_level(const nlod_lod_grid* grid, uint32_t lod, const uint32_t** d_table_out,
                        nlod_voxinfo* info_out);
/* First occupied voxel of level `lod` along each ray o + t d, t >= t_min
 * (d_t_min nullable: 0), skipping voxel d_prev_voxel[i] (nullable; linear
 * index x + y R + z R^2 or 0xffffffff): d_voxel_out = its linear index or
 * 0xffffffff when the ray leaves the grid, d_t_near_out / d_t_far_out the
 * ray's entry and exit parameters of that voxel's box (-1 when none).
 * Replaces first_active_voxel_intersection (neural_lod_learning_common.cuh:
 * 784-832) as raymarch_impl uses it (wavefront_neural_pt.cuh:89-141).
 * Origins and directions float[n][3], world space.  Asynchronous. */
int nlod_lod_grid_march(const nlod_lod_grid* grid, void* stream, uint32_t lod, uint32_t n, const float* d_origins,
                        const float* d_dirs, const float* d_t_min, const uint32_t* d_prev_voxel,
                        uint32_t* d_voxel_out, float* d_t_near_out, float* d_t_far_out);
/* Number of occupied voxels of level `lod` (the reference's
 * gridMetaData()->activeVoxelCount(), neural_lod_learning_init.cu:272).
 * Needs a power-of-two level resolution <= 1024.  Synchronizes the stream. */
int nlod_lod_grid_active_voxels(const nlod_lod_grid* grid, void* stream, uint32_t lod, uint64_t* count_out);
/* Per-voxel visibility thresholds of level `lod`: d_thresholds float[n], one
 * per occupied voxel in increasing Morton order of (x, y, z) (x in bit 0) --
 * the order and file layout of the reference's optimal-threshold files
 * (read_vector_from_file, neural_lod_learning.h:482-496) -- each stored as
 * min(max(t, min_threshold), max_threshold).  n must equal the level's
 * occupied voxel count.  Replaces update_grid_with_thresholds
 * (neural_lod_learning.cu:50-121).  Levels without thresholds report 0.5 for
 * every voxel (the initial value, neural_lod_learning_init.cu:242).  Needs a
 * power-of-two level resolution <= 1024; keeps R^3 floats on the device.
 * Synchronizes the stream. */
int nlod_lod_grid_set_thresholds(nlod_lod_grid* grid, void* stream, uint32_t lod, uint64_t n,
                                 const float* d_thresholds, float min_threshold, float max_threshold);
/* nlod_lod_grid_march that also writes the hit voxel's threshold,
 * d_threshold_out float[n] (-1 where the ray leaves the grid), as
 * first_active_voxel_intersection returns voxel_threshold
 * (neural_lod_learning_common.cuh:784-832). */
int nlod_lod_grid_march_thresholds(const nlod_lod_grid* grid, void* stream, uint32_t lod, uint32_t n,
                                   const float* d_origins, const float* d_dirs, const float* d_t_min,
                                   const uint32_t* d_prev_voxel, uint32_t* d_voxel_out, float* d_t_near_out,
                                   float* d_t_far_out, float* d_threshold_out);

/* ------------------------------------------------ neural LoD path tracer */

/* Pinhole camera of the reference's ViewParams (rptr/vulkan/gpu_params.glsl:
 * 70-74): pixel (x, y) of a width x height framebuffer looks along
 * normalize((x + u) / width * du + (y + v) / height * dv + dir_top_left) from
 * position, (u, v) a per-pixel jitter (camera_raygen,
 * rptr/cuda/wavefront_neural_pt.cu:9-49).  World space. */
typedef struct nlod_camera {
    float position[3];
    float dir_top_left[3];
    float du[3];
    float dv[3];
    uint32_t width, height;
} nlod_camera;

enum nlod_tracer_schedule {
    NLOD_TRACER_AUTO = 0,       /* persistent where it applies, else the reference's */
    NLOD_TRACER_REFERENCE = 1,  /* the reference's wavefront rounds, one host count readback per round */
    NLOD_TRACER_PERSISTENT = 2  /* FP16: one launch per frame, device ray queue, no host readback */
};

/* WavefrontNeuralThroughputVisibilityLod parameters
 * (rptr/cuda/neural-lod/wavefront_neural_throughput_visibility_lod.h:19-31);
 * nlod_tracer_default_params fills the reference's defaults. */
typedef struct nlod_tracer_params {
    uint32_t lod;                        /* pyramid level marched (current_lod) */
    uint32_t max_depth;                  /* bounces (3) */
    float background_color[3];           /* paths escaping at bounce 0 (1, 1, 1) */
    float envmap_color[3];               /* constant environment, later bounces (1, 1, 1) */
    float inference_min_extent_dilation; /* 1e-4 */
    float voxel_extent_dilation_outward; /* dilation the visibility network was trained with (1e-4) */
    uint32_t apply_rr, rr_start_bounce;  /* Russian roulette on the throughput (off, 1) */
    uint32_t apply_visibility_rr;        /* roulette on unoccluded visibility rounds >= 5 (off) */
    uint32_t stochastic_threshold;       /* occluded when visibility >= a uniform draw (off) */
    uint32_t use_hg;                     /* Henyey-Greenstein directions for the throughput query (off) */
    float hg_g;                          /* 0.5 */
    uint32_t throughput_log_learning;    /* outputs are log(1 + throughput) (off) */
    uint32_t schedule;                   /* enum nlod_tracer_schedule */
} nlod_tracer_params;

typedef struct nlod_tracer nlod_tracer;

int nlod_tracer_default_params(nlod_tracer_params* params_out);

/* Path tracer over one level of a LoD pyramid with a visibility network of 6
 * inputs (the segment's entry and exit points in the voxel, normalized to the
 * box) and a throughput network of 9 inputs (voxel centre, (wo + 1) / 2,
 * (wi + 1) / 2) or 10 (the level appended) and 3 outputs.  The grid and the
 * models must outlive the tracer; the networks' inference parameters are used.
 * Replaces the WavefrontNeuralThroughputVisibilityLod render extension
 * (wavefront_neural_throughput_visibility_lod.cu:420-462, 727-794). */
int nlod_tracer_create(const nlod_lod_grid* grid, nlod_model* visibility, nlod_model* throughput,
                       nlod_tracer** tracer_out);
int nlod_tracer_destroy(nlod_tracer* tracer);

/* One sample per pixel: voxel_tracing(sample_idx, ...)
 * (wavefront_neural_throughput_visibility_lod.cu:464-720) with random
 * numbers from get_lcg_rng(sample_idx, sample_offset, pixel).  Per-voxel
 * thresholds come from nlod_lod_grid_set_thresholds (0.5 where unset).
 * d_illumination float[width * height][3] gets each path's contribution
 * (what accumulate_neural_bounce writes), d_inferences (nullable)
 * uint32[width * height] its visibility-network queries (needed_inferences),
 * d_bounces (nullable) the bounce at which it ended.  The reference schedule
 * synchronizes the stream once per round; the persistent one is asynchronous,
 * allocates nothing after its first call and can be captured in a hipGraph. */
int nlod_tracer_render(nlod_tracer* tracer, void* stream, const nlod_camera* camera,
                       const nlod_tracer_params* params, uint32_t sample_idx, uint32_t sample_offset,
                       float* d_illumination, uint32_t* d_inferences, uint32_t* d_bounces);
/* Counts of the last persistent-schedule frame: stats_out[0] visibility
 * network evaluations (one per 16-slot tile of a wave), [1] the visibility
 * queries they served, [2] and [3] the same for the throughput network, [4]
 * the waves' rounds.  [1] / (16 [0]) is the tile occupancy of the network
 * evaluations.  Synchronizes the stream; zeros before the first persistent
 * frame.  No counterpart in the reference. */
int nlod_tracer_stats(nlod_tracer* tracer, void* stream, uint64_t stats_out[5]);
/* The schedule nlod_tracer_render would use for these parameters. */
int nlod_tracer_schedule(const nlod_tracer* tracer, const nlod_tracer_params* params, int* schedule_out);

/* ------------------------------------------ per-voxel visibility thresholds */

/* Scores of NeuralLodLearning::ThresholdOptimMetrics, same order
 * (neural_lod_learning.h:105-118; formulas neural_lod_learning_common.cuh:
 * 836-927). */
enum nlod_threshold_metric {
    NLOD_METRIC_PRECISION = 0,
    NLOD_METRIC_RECALL = 1,
    NLOD_METRIC_FSCORE = 2,
    NLOD_METRIC_FSCORE_MACRO = 3,
    NLOD_METRIC_FSCORE_WEIGHTED = 4,
    NLOD_METRIC_OPTIMAL_FSCORE_WEIGHTED = 5,
    NLOD_METRIC_MCC = 6,
    NLOD_METRIC_COHEN_KAPPA = 7,
    NLOD_METRIC_ACCURACY = 8,
    NLOD_METRIC_BALANCED_ACCURACY = 9,
    NLOD_METRIC_FOWLKES_MALLOWS = 10,
    NLOD_METRIC_YOUDEN_J = 11
};

/* Per voxel v: the threshold t / (T - 1) maximising `metric` of the
 * predictions p > t / (T - 1) against labels (1.0 = occluded) over samples
 * [v S, v S + S) of d_labels / d_probas, ties resolved to the middle of the
 * plateau that follows the first maximum.  Replaces compute_confusion_matrices,
 * compute_scores and compute_optimal_thresholds (neural_lod_learning.cu:
 * 122-266).  d_thresholds_out float[n_voxels]; d_scores_out (nullable)
 * float[n_voxels][T].  T in [2, 1024].  Asynchronous. */
int nlod_optimal_thresholds(void* stream, uint32_t n_voxels, uint32_t samples_per_voxel, const float* d_labels,
                            const float* d_probas, uint32_t n_thresholds, int metric, float beta,
                            float* d_thresholds_out, float* d_scores_out);

#ifdef __cplusplus
}
#endif

#endif /* NLOD_H */
