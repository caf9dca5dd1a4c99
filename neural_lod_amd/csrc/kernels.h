// Host-side launchers of the HIP kernels (one translation unit per family).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.h"
#include "nlod.h"

namespace nlod {

// ---- encoding.hip ----------------------------------------------------------
void fill_f32(hipStream_t stream, float* p, size_t n, float v);
// inputs AoS [N][n_input_dims] -> encoded SoA float [enc_width][N]
void encode_forward_f32(hipStream_t stream, const Layout& L, const float* params, const float* in, uint32_t N,
                        float* enc);
// accumulate grid parameter gradients from dL/d(enc) (SoA) into grads (full parameter vector)
void encode_backward_params(hipStream_t stream, const Layout& L, const float* in, uint32_t N, const float* d_enc,
                            float* grads);
bool encode_backward_input_supported(const Layout& L);
// dL/dinput AoS [N][n_input_dims] (SH and identity segments)
void encode_backward_input(hipStream_t stream, const Layout& L, const float* in, uint32_t N, const float* d_enc,
                           float* d_in);

// ---- mlp_fp32.hip (plain FP32 form) -------------------------------------------
// enc SoA [K][N] -> out AoS [N][n_output]; hidden (nullable) SoA [L][W][N] post-activation
void mlp_forward_f32(hipStream_t stream, const NetShape& net, const float* weights, const float* enc, uint32_t N,
                     float* out, float* hidden);
// dL/dout AoS [N][n_output] -> delta (SoA [L][W][N], pre-activation gradients), delta_out SoA [16][N],
// dL/denc SoA [K][N] (nullable)
void mlp_backward_delta_f32(hipStream_t stream, const NetShape& net, const float* weights, const float* hidden,
                            const float* out, const float* d_out, uint32_t N, float* delta, float* delta_out,
                            float* d_enc);
// weight gradients dW = delta . act^T summed over the batch (FP32), written (or added) into grads
void mlp_weight_grad_f32(hipStream_t stream, const NetShape& net, const float* enc, const float* hidden,
                         const float* delta, const float* delta_out, uint32_t N, float* grads, float* workspace,
                         size_t workspace_floats, bool accumulate);
size_t mlp_weight_grad_workspace_floats(const NetShape& net, uint32_t N);

// ---- voxelize.hip -------------------------------------------------------------
void voxinfo_from_bbox(const float bmin[3], const float bmax[3], uint32_t G, uint32_t flags, uint64_t n_tri,
                       struct nlod_voxinfo* info);
// no_alloc: the tiled path takes its scratch from the library's fixed device
// array instead of the per-stream buffer (fails when the job needs more)
void voxelize(hipStream_t stream, const struct nlod_voxinfo* info, const float* d_tris, uint32_t* d_table,
              uint32_t flags, bool no_alloc = false);
void mesh_bbox(hipStream_t stream, const float* d_tris, uint64_t n_tri, float bmin[3], float bmax[3]);

}  // namespace nlod
