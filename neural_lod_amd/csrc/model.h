// NetworkWithInputEncoding on gfx950: the model object behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "grid_backward.h"
#include "kernels.h"
#include "layout.h"
#include "nlod_common.h"

namespace nlod {

enum Precision : int { PREC_FP32 = 0, PREC_FP16 = 1 };
enum GradMode : int { GRAD_IGNORE = 0, GRAD_OVERWRITE = 1, GRAD_ACCUMULATE = 2 };

struct Model;

// Saved state of one forward pass (tcnn's ForwardContext).
struct ForwardCtx {
    uint32_t N = 0;
    bool use_inference_params = false;
    bool prepare_input_gradients = false;
    bool recompute = false;  // activations were not stored (FP16 path, large batches)
    DevBuf enc;     // FP32: encoded SoA float [K][N]; FP16: encoded half [N/16 tiles]...
    DevBuf hidden;  // FP32: SoA float [L][W][N]; FP16: half tiles
    DevBuf out;     // post-activation outputs float [N][n_output]
};

struct Model {
    Layout L;
    int precision = PREC_FP32;
    // Incremented when a buffer of the model (or of its grid backward) is freed
    // or moved, and when the parameter binding changes: a hipGraph that captured
    // the model's addresses is valid only while the epoch it was stamped with holds.
    uint64_t epoch = 0;
    const float* params = nullptr;
    const float* params_inference = nullptr;
    bool inference_aliased = true;  // params_inference == params: one set of FP16 copies
    // FP16 form: MLP weights packed for the MFMA kernels (forward and transposed for
    // backward), the device copy of the encoding description, and FP16 copies of
    // the whole parameter vector (the grid tables the encoding pass gathers)
    DevBuf packed, packed_bwd, table16;
    DevBuf packed_inference, packed_bwd_inference, table16_inference;
    DevBuf program;
    DevBuf ws_feat;
    DevBuf ws_pos;  // [3][N] grid positions of the FP16 encoding pass
    void* pos_scratch(uint32_t N);  // ws_pos sized for N queries, or null when unused
    // scratch
    DevBuf ws_enc, ws_delta, ws_delta_out, ws_denc, ws_wgrad;
    // recomputed activations of one chunk (batches above kRecomputeAbove)
    DevBuf ws_hidden, ws_out;
    // hash-grid parameter gradients (binned, no global atomics)
    std::unique_ptr<GridBackward> grid_bwd;

    Model(uint32_t n_in, uint32_t n_out, const Json& enc, const Json& net, int prec);

    void set_params(hipStream_t s, const float* p, const float* p_inf);
    // FP16 form: repack the MLP weights after the FP32 parameters changed on the
    // device (the FP16 table copies are written by whoever changed them)
    void refresh_packed(hipStream_t s, bool training, bool inference);
    void require_params(bool inference) const;
    const float* weights(bool inference) const { return inference ? params_inference : params; }
    bool use_inference_copy(bool inference) const { return inference && !inference_aliased; }
    const void* packed_fwd(bool inf) const { return use_inference_copy(inf) ? packed_inference.p : packed.p; }
    const void* packed_back(bool inf) const { return use_inference_copy(inf) ? packed_bwd_inference.p : packed_bwd.p; }
    void* table_half(bool inf) const { return use_inference_copy(inf) ? table16_inference.p : table16.p; }

    void encode(hipStream_t s, uint32_t N, const float* in, float* enc_soa, bool inference);
    void inference(hipStream_t s, uint32_t N, const float* in, float* out);
    std::unique_ptr<ForwardCtx> forward(hipStream_t s, uint32_t N, const float* in, float* out, bool use_inference,
                                        bool prepare_input_gradients);
    // forward into an existing context (its buffers only grow)
    void forward_into(hipStream_t s, uint32_t N, const float* in, float* out, bool use_inference,
                      bool prepare_input_gradients, ForwardCtx& ctx);
    // grid_grads_zero: the grid part of d_params is already zero (skips its clearing under OVERWRITE)
    void backward(hipStream_t s, ForwardCtx& ctx, const float* in, const float* d_out, float* d_params,
                  float* d_in, int grad_mode, bool grid_grads_zero = false);

  private:
    // outputs only (no saved activations), chunked
    void forward_outputs(hipStream_t s, uint32_t N, const float* in, float* out, bool use_inference);
    // activations of a batch into enc / hidden / out
    void forward_saved(hipStream_t s, uint32_t N, const float* in, bool use_inference, void* enc, void* hidden,
                       float* out);
    // backward of one batch from its saved activations; net_accumulate: add the
    // network gradients to d_params instead of overwriting them
    void backward_saved(hipStream_t s, uint32_t N, bool use_inference, const float* in, const float* d_out,
                        const void* enc, const void* hidden, const float* out, float* d_params, float* d_in,
                        bool param_grads, bool net_accumulate, bool grid_overwrite);

  public:
    void init_params_host(uint32_t seed, float* out) const;
    // grid gradients from dL/d(enc) SoA [K][N]: overwrite stores every table entry
    // (untouched ones get 0), otherwise they are added
    void grid_backward(hipStream_t s, const float* in, uint32_t N, const float* d_enc, float* d_params,
                       bool overwrite);
    // true when grid_backward(overwrite) writes every table entry itself
    bool grid_overwrite_covers_tables() const { return grid_bwd && grid_bwd->supported(); }
    // FP16 form: one training pass (encoding, fused MLP forward + loss + backward +
    // dW, grid backward) writing every gradient (overwrite) and the FP64 loss
    // partials; returns the number of loss partials written (0: not available
    // for this layout, nothing was done)
    uint32_t train_fused(hipStream_t s, uint32_t N, const float* in, const float* target, int loss_type,
                         const float* loss_scale, double* loss_partial, uint32_t max_partials, float* d_params);
};

// Batches above this size recompute their activations chunk by chunk in the
// backward pass instead of storing them in the forward pass.
constexpr uint32_t kRecomputeAbove = 1u << 20;

// ---- mlp_mfma.hip (FP16 MFMA form) ----------------------------------------------
bool mfma_supported(const Layout& L);
size_t mfma_packed_bytes(const NetShape& net);
size_t mfma_program_bytes();
size_t mfma_fused_smem_bytes(const NetShape& net);
// encoding description for the fused kernels, copied to device memory (synchronous)
void mfma_upload_program(const Layout& L, void* d_program);
// FP16 weight fragments of the forward packing and (packed_bwd non-null) the
// backward packing of the transposed weights, in one launch
void mfma_pack_weights(hipStream_t stream, const NetShape& net, const float* weights, void* packed,
                       void* packed_bwd);
// FP32 -> FP16 copy (the grid tables read by the encoding pass)
void mfma_to_half(hipStream_t stream, const float* src, void* dst, size_t n);
// bytes of the FP16 feature buffer for N queries ([N/16][K/8][16][8] halves)
size_t mfma_feature_bytes(const Layout& L, uint32_t N);
// encoding pass (per-level grid gathers + non-grid rows) -> FP16 feature fragments
// pos_scratch (nullable): mfma_pos_scratch_bytes(L, N) bytes of device memory
// for a [3][N] copy of the grid positions; 0 bytes means the layout does not
// use one.
void mfma_encode(hipStream_t stream, const Layout& L, const void* d_program, const void* table16, const float* in,
                 uint32_t N, void* feat, void* pos_scratch = nullptr);
size_t mfma_pos_scratch_bytes(const Layout& L, uint32_t N);
// MLP on MFMA from the feature fragments -> out float[N][n_output]
void mfma_mlp_forward(hipStream_t stream, const Layout& L, const void* packed, const void* feat, uint32_t N,
                      float* out);


// ---- mlp_mfma_train.hip (FP16 MFMA training kernels) ------------------------------
size_t mfma_packed_bwd_bytes(const NetShape& net);
// saved hidden activations: half [L][N/16][W][16]
size_t mfma_hidden_bytes(const Layout& L, uint32_t N);
// pre-activation gradients: half [L][N/16][W][16] + output layer [N/16][16][16]
size_t mfma_delta_bytes(const Layout& L, uint32_t N);
// forward storing hidden activations and the outputs float[N][n_output]
void mfma_mlp_forward_train(hipStream_t stream, const Layout& L, const void* packed, const void* feat, uint32_t N,
                            void* hidden, float* out);
// deltas of every layer from dL/dout float[N][n_output]; dL/denc SoA float [K][N] (nullable)
void mfma_mlp_backward(hipStream_t stream, const Layout& L, const void* packed_bwd, const void* hidden,
                       const float* y, const float* dy, uint32_t N, void* delta, float* d_enc);
size_t mfma_wgrad_workspace_floats(const Layout& L, uint32_t N);
// network weight gradients (FP32, deterministic) written or added into grads[0 .. n_network_params)
void mfma_weight_grad(hipStream_t stream, const Layout& L, const void* feat, const void* hidden, const void* delta,
                      uint32_t N, float* grads, float* workspace, size_t workspace_floats, bool accumulate);


// ---- mlp_mfma_fused_train.hip (forward + loss + backward + dW in one kernel) -------
bool mfma_fused_train_supported(const Layout& L);
// blocks of the fused kernel = loss partials and dW partial rows it writes
uint32_t mfma_fused_train_blocks(const Layout& L, uint32_t N);
// loss_partial[blocks] (FP64), d_enc SoA float [K][N] (nullable), w_partial [blocks][n_network_params]
void mfma_fused_train(hipStream_t stream, const Layout& L, const void* packed, const void* packed_bwd,
                      const void* feat, const float* target, uint32_t N, int loss_type, const float* loss_scale,
                      double* loss_partial, float* d_enc, float* w_partial);
// out[e] (+)= sum over rows of partial[rows][n], in row order
void partial_reduce(hipStream_t stream, uint32_t n, uint32_t rows, const float* partial, float* out, bool accumulate);

}  // namespace nlod
