// Input encodings of the neural-LoD networks (gfx950):
//   multiresolution hash grid forward / parameter backward (tiny-cuda-nn
//   encodings/grid.h:220-517), spherical harmonics forward / input backward
//   (encodings/spherical_harmonics.h:46-390), identity (encodings/identity.h:45-68),
//   and the padding rows of a composite encoding.
// Encoded features are SoA: row r of the encoded vector for query i is
// enc[r * N + i], as the reference stores encoding outputs.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "kernels.h"
#include "nlod_common.h"
#include "sh_basis.h"

namespace nlod {
namespace {

template <int F>
struct FVec;
template <>
struct FVec<1> { float v[1]; };
template <>
struct FVec<2> { float v[2]; };
template <>
struct FVec<4> { float v[4]; };
template <>
struct FVec<8> { float v[8]; };

template <int F>
__device__ __forceinline__ FVec<F> load_entry(const float* __restrict__ p) {
    FVec<F> r;
    if constexpr (F == 1) {
        r.v[0] = p[0];
    } else if constexpr (F == 2) {
        float2 a = *reinterpret_cast<const float2*>(p);
        r.v[0] = a.x; r.v[1] = a.y;
    } else {
#pragma unroll
        for (int k = 0; k < F; k += 4) {
            float4 a = *reinterpret_cast<const float4*>(p + k);
            r.v[k] = a.x; r.v[k + 1] = a.y; r.v[k + 2] = a.z; r.v[k + 3] = a.w;
        }
    }
    return r;
}

// One thread per (query, level): blockIdx.y is the level, so each level's
// table slice is reused by a whole wave of blocks (grid.h:1053-1058).
template <int F>
__global__ __launch_bounds__(256) void grid_fwd_kernel(uint32_t N, GridArgs a, const float* __restrict__ in,
                                                       const float* __restrict__ table, float* __restrict__ enc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t level = blockIdx.y;
    const float* x = in + (size_t)i * a.in_stride + a.in_offset;
    const float scale = a.lv.scale[level];
    const uint32_t res = a.lv.res[level];
    const uint32_t off = a.lv.offset[level];
    const uint32_t hs = a.lv.offset[level + 1] - off;
    const float* tab = table + (size_t)off * F;
    float fr[3];
    uint32_t cell[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) grid_pos(x[d], scale, a.lv.interp, fr[d], cell[d]);
    float acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.0f;
    if (a.lv.interp == INTERP_NEAREST) {
        const uint32_t idx = grid_index(a.lv.grid_type, hs, res, cell[0], cell[1], cell[2]);
        FVec<F> v = load_entry<F>(tab + (size_t)idx * F);
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = v.v[f];
    } else {
        FVec<F> vals[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const uint32_t idx = grid_index(a.lv.grid_type, hs, res, cell[0] + (c & 1), cell[1] + ((c >> 1) & 1),
                                            cell[2] + ((c >> 2) & 1));
            vals[c] = load_entry<F>(tab + (size_t)idx * F);
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            float w = 1.0f;
#pragma unroll
            for (int d = 0; d < 3; ++d) w *= ((c >> d) & 1) ? fr[d] : 1.0f - fr[d];
#pragma unroll
            for (int f = 0; f < F; ++f) acc[f] += w * vals[c].v[f];
        }
    }
#pragma unroll
    for (int f = 0; f < F; ++f) enc[(size_t)(a.out_offset + level * F + f) * N + i] = acc[f];
}

// Scatter dL/dy * weight into the 8 corners (grid.h:394-517), FP32 atomics.
template <int F>
__global__ __launch_bounds__(256) void grid_bwd_kernel(uint32_t N, GridArgs a, const float* __restrict__ in,
                                                       const float* __restrict__ d_enc, float* __restrict__ grad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t level = blockIdx.y;
    float g[F];
    bool any = false;
#pragma unroll
    for (int f = 0; f < F; ++f) {
        g[f] = d_enc[(size_t)(a.out_offset + level * F + f) * N + i];
        any |= (g[f] != 0.0f);
    }
    if (!any) return;  // adding zeros changes nothing
    const float* x = in + (size_t)i * a.in_stride + a.in_offset;
    const float scale = a.lv.scale[level];
    const uint32_t res = a.lv.res[level];
    const uint32_t off = a.lv.offset[level];
    const uint32_t hs = a.lv.offset[level + 1] - off;
    float* gt = grad + (size_t)off * F;
    float fr[3];
    uint32_t cell[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) grid_pos(x[d], scale, a.lv.interp, fr[d], cell[d]);
    if (a.lv.interp == INTERP_NEAREST) {
        const uint32_t idx = grid_index(a.lv.grid_type, hs, res, cell[0], cell[1], cell[2]);
#pragma unroll
        for (int f = 0; f < F; ++f) unsafeAtomicAdd(gt + (size_t)idx * F + f, g[f]);
        return;
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float w = 1.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) w *= ((c >> d) & 1) ? fr[d] : 1.0f - fr[d];
        const uint32_t idx = grid_index(a.lv.grid_type, hs, res, cell[0] + (c & 1), cell[1] + ((c >> 1) & 1),
                                        cell[2] + ((c >> 2) & 1));
#pragma unroll
        for (int f = 0; f < F; ++f) unsafeAtomicAdd(gt + (size_t)idx * F + f, w * g[f]);
    }
}

struct ShArgs {
    uint32_t in_offset, in_stride, out_offset, n_pad, degree;
};

__device__ __forceinline__ void sh_powers(const float* x, float* X, float* Y, float* Z) {
    const float vx = x[0] * 2.f - 1.f, vy = x[1] * 2.f - 1.f, vz = x[2] * 2.f - 1.f;
    X[0] = Y[0] = Z[0] = 1.0f;
#pragma unroll
    for (int k = 1; k < 8; ++k) {
        X[k] = X[k - 1] * vx;
        Y[k] = Y[k - 1] * vy;
        Z[k] = Z[k - 1] * vz;
    }
}

__global__ __launch_bounds__(256) void sh_fwd_kernel(uint32_t N, ShArgs a, const float* __restrict__ in,
                                                     float* __restrict__ enc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float X[8], Y[8], Z[8];
    sh_powers(in + (size_t)i * a.in_stride + a.in_offset, X, Y, Z);
    float* o = enc + (size_t)a.out_offset * N + i;
    for (uint32_t p = 0; p < a.n_pad; ++p) o[(size_t)p * N] = 1.0f;  // pad rows first
    o += (size_t)a.n_pad * N;
    sh_basis(a.degree, X, Y, Z, [&](int k, float v) { o[(size_t)k * N] = v; });
}

// dL/dinput of the SH encoding (spherical_harmonics.h:153-390): chain rule
// through (2*in - 1), accumulated into d_in[i][in_offset + d].
__global__ __launch_bounds__(256) void sh_bwd_input_kernel(uint32_t N, ShArgs a, const float* __restrict__ in,
                                                           const float* __restrict__ d_enc, float* __restrict__ d_in,
                                                           uint32_t d_in_stride) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float X[8], Y[8], Z[8];
    sh_powers(in + (size_t)i * a.in_stride + a.in_offset, X, Y, Z);
    const float* g = d_enc + (size_t)(a.out_offset + a.n_pad) * N + i;
    float gx = 0.f, gy = 0.f, gz = 0.f;
    sh_basis_dx(a.degree, X, Y, Z, [&](int k, float v) { gx += v * g[(size_t)k * N]; });
    sh_basis_dy(a.degree, X, Y, Z, [&](int k, float v) { gy += v * g[(size_t)k * N]; });
    sh_basis_dz(a.degree, X, Y, Z, [&](int k, float v) { gz += v * g[(size_t)k * N]; });
    float* o = d_in + (size_t)i * d_in_stride + a.in_offset;
    o[0] += 2.0f * gx;
    o[1] += 2.0f * gy;
    o[2] += 2.0f * gz;
}

struct IdArgs {
    uint32_t in_offset, in_stride, out_offset, n_out, n_pad;
    float scale, offset;
};

__global__ __launch_bounds__(256) void identity_fwd_kernel(uint32_t N, IdArgs a, const float* __restrict__ in,
                                                           float* __restrict__ enc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float* x = in + (size_t)i * a.in_stride + a.in_offset;
    for (uint32_t k = 0; k < a.n_out; ++k) enc[(size_t)(a.out_offset + k) * N + i] = x[k] * a.scale + a.offset;
    for (uint32_t p = 0; p < a.n_pad; ++p) enc[(size_t)(a.out_offset + a.n_out + p) * N + i] = 1.0f;
}

__global__ __launch_bounds__(256) void identity_bwd_input_kernel(uint32_t N, IdArgs a, const float* __restrict__ d_enc,
                                                                 float* __restrict__ d_in, uint32_t d_in_stride) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (uint32_t k = 0; k < a.n_out; ++k)
        d_in[(size_t)i * d_in_stride + a.in_offset + k] += a.scale * d_enc[(size_t)(a.out_offset + k) * N + i];
}

__global__ __launch_bounds__(256) void fill_rows_kernel(float* __restrict__ p, size_t n, float v) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = v;
}

GridArgs make_grid_args(const Segment& s, uint32_t in_stride) {
    GridArgs a{};
    a.lv = s.levels;
    a.in_offset = s.in_offset;
    a.in_stride = in_stride;
    a.out_offset = s.out_offset;
    return a;
}

template <typename K>
void launch_by_F(uint32_t F, K&& k) {
    switch (F) {
        case 1: k(std::integral_constant<int, 1>{}); break;
        case 2: k(std::integral_constant<int, 2>{}); break;
        case 4: k(std::integral_constant<int, 4>{}); break;
        case 8: k(std::integral_constant<int, 8>{}); break;
        default: fail(NLOD_ERR_CONFIG, "unsupported n_features_per_level");
    }
}

}  // namespace

void fill_f32(hipStream_t stream, float* p, size_t n, float v) {
    if (n == 0) return;
    const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(fill_rows_kernel, dim3(blocks), dim3(256), 0, stream, p, n, v);
    NLOD_LAUNCH_CHECK();
}

void encode_forward_f32(hipStream_t stream, const Layout& L, const float* params, const float* in, uint32_t N,
                        float* enc) {
    if (N == 0) return;
    const uint32_t nb = div_round_up(N, 256);
    for (const Segment& s : L.segs) {
        if (s.type == SEG_GRID) {
            GridArgs a = make_grid_args(s, L.n_input_dims);
            const float* table = params + s.param_offset;
            launch_by_F(s.F, [&](auto Fc) {
                hipLaunchKernelGGL(grid_fwd_kernel<decltype(Fc)::value>, dim3(nb, s.n_levels), dim3(256), 0, stream, N,
                                   a, in, table, enc);
            });
            NLOD_LAUNCH_CHECK();
            if (s.n_pad) fill_f32(stream, enc + (size_t)(s.out_offset + s.n_out) * N, (size_t)s.n_pad * N, 0.0f);
        } else if (s.type == SEG_SH) {
            ShArgs a{s.in_offset, L.n_input_dims, s.out_offset, s.n_pad, s.degree};
            hipLaunchKernelGGL(sh_fwd_kernel, dim3(nb), dim3(256), 0, stream, N, a, in, enc);
            NLOD_LAUNCH_CHECK();
        } else {
            IdArgs a{s.in_offset, L.n_input_dims, s.out_offset, s.n_out, s.n_pad, s.id_scale, s.id_offset};
            hipLaunchKernelGGL(identity_fwd_kernel, dim3(nb), dim3(256), 0, stream, N, a, in, enc);
            NLOD_LAUNCH_CHECK();
        }
    }
}

void encode_backward_params(hipStream_t stream, const Layout& L, const float* in, uint32_t N, const float* d_enc,
                            float* grads) {
    if (N == 0) return;
    const uint32_t nb = div_round_up(N, 256);
    for (const Segment& s : L.segs) {
        if (s.type != SEG_GRID) continue;
        GridArgs a = make_grid_args(s, L.n_input_dims);
        float* g = grads + s.param_offset;
        launch_by_F(s.F, [&](auto Fc) {
            hipLaunchKernelGGL(grid_bwd_kernel<decltype(Fc)::value>, dim3(nb, s.n_levels), dim3(256), 0, stream, N, a,
                               in, d_enc, g);
        });
        NLOD_LAUNCH_CHECK();
    }
}

bool encode_backward_input_supported(const Layout& L) {
    for (const Segment& s : L.segs)
        if (s.type == SEG_GRID) return false;
    return true;
}

void encode_backward_input(hipStream_t stream, const Layout& L, const float* in, uint32_t N, const float* d_enc,
                           float* d_in) {
    if (N == 0) return;
    const uint32_t nb = div_round_up(N, 256);
    fill_f32(stream, d_in, (size_t)N * L.n_input_dims, 0.0f);
    for (const Segment& s : L.segs) {
        if (s.type == SEG_SH) {
            ShArgs a{s.in_offset, L.n_input_dims, s.out_offset, s.n_pad, s.degree};
            hipLaunchKernelGGL(sh_bwd_input_kernel, dim3(nb), dim3(256), 0, stream, N, a, in, d_enc, d_in,
                               L.n_input_dims);
        } else if (s.type == SEG_IDENTITY) {
            IdArgs a{s.in_offset, L.n_input_dims, s.out_offset, s.n_out, s.n_pad, s.id_scale, s.id_offset};
            hipLaunchKernelGGL(identity_bwd_input_kernel, dim3(nb), dim3(256), 0, stream, N, a, d_enc, d_in,
                               L.n_input_dims);
        } else {
            fail(NLOD_ERR_UNSUPPORTED, "input gradients through a grid encoding are not supported");
        }
        NLOD_LAUNCH_CHECK();
    }
}

}  // namespace nlod
