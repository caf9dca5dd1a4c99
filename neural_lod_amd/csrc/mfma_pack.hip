// FP16 weight packing and the device copy of the encoding program for the
// MFMA kernels (layouts documented in mfma_common.h).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>

#include "mfma_common.h"
#include "model.h"

namespace nlod {
namespace {

// Elements [0, total_fwd) of the forward packing, then (out_bwd non-null) the
// backward packing, in one launch.
__global__ void pack_kernel(uint32_t W, uint32_t K, uint32_t Kp, uint32_t L, const float* __restrict__ w,
                            half_t* __restrict__ out_fwd, uint32_t total_fwd, half_t* __restrict__ out_bwd,
                            uint32_t total_bwd) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < total_fwd) out_fwd[e] = (half_t)packed_fwd_value(W, K, Kp, L, w, e);
    else if (out_bwd && e - total_fwd < total_bwd) out_bwd[e - total_fwd] = (half_t)packed_bwd_value(W, K, L, w, e - total_fwd);
}

EncProgram build_program(const Layout& L) {
    EncProgram P{};
    P.n_segs = (uint32_t)L.segs.size();
    P.n_input_dims = L.n_input_dims;
    P.enc_width = L.enc_width;
    P.k_padded = k_padded_of(L.net);
    for (size_t i = 0; i < L.segs.size(); ++i) {
        const Segment& s = L.segs[i];
        DevSegment& d = P.segs[i];
        d.type = s.type;
        d.in_offset = s.in_offset;
        d.out_offset = s.out_offset;
        d.n_out = s.n_out;
        d.n_pad = s.n_pad;
        d.degree = s.degree;
        d.id_scale = s.id_scale;
        d.id_offset = s.id_offset;
        d.param_offset = s.param_offset;
        d.lv = s.levels;
        if (s.type == SEG_GRID)
            for (uint32_t r = s.out_offset; r < s.out_offset + s.n_out; ++r) P.grid_rows[r / 8] |= (uint8_t)(1u << (r % 8));
    }
    return P;
}

}  // namespace

int cu_count() {
    static int n = [] {
        int dev = 0, c = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) return 256;
        return c;
    }();
    return n;
}

int occupancy_blocks(const void* kernel, int block, size_t smem) {
    static std::mutex mu;
    static std::map<std::tuple<int, const void*, int, size_t>, int> cache;
    int dev = 0;
    NLOD_HIP(hipGetDevice(&dev));
    const auto key = std::make_tuple(dev, kernel, block, smem);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    if (smem) NLOD_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    int b = 0;
    NLOD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, block, smem));
    b = b > 0 ? b : 1;
    cache.emplace(key, b);
    return b;
}

size_t mfma_packed_bytes(const NetShape& net) { return packed_halves(net) * sizeof(half_t); }

size_t mfma_program_bytes() { return sizeof(EncProgram); }

void mfma_upload_program(const Layout& L, void* d_program) {
    const EncProgram host = build_program(L);
    NLOD_HIP(hipMemcpy(d_program, &host, sizeof(EncProgram), hipMemcpyHostToDevice));
}

void mfma_pack_weights(hipStream_t stream, const NetShape& net, const float* weights, void* packed,
                       void* packed_bwd) {
    const uint32_t total_fwd = (uint32_t)packed_halves(net);
    const uint32_t total_bwd = packed_bwd ? (uint32_t)packed_bwd_halves(net) : 0u;
    hipLaunchKernelGGL(pack_kernel, dim3(div_round_up(total_fwd + total_bwd, 256)), dim3(256), 0, stream, net.width,
                       net.n_input, k_padded_of(net), net.n_hidden_layers, weights, static_cast<half_t*>(packed),
                       total_fwd, static_cast<half_t*>(packed_bwd), total_bwd);
    NLOD_LAUNCH_CHECK();
}

}  // namespace nlod
