// One FP16 network evaluated by a single wave on a 16-query tile, with its
// packed weights in LDS: the building block of the kernels that run both LoD
// networks in one launch (the wavefront engine and the path tracer).
//
// Lane (g, n) = (lane / 16, lane % 16) works on query n of the tile; the four
// lane groups encode the four 8-feature windows of each 32-wide k-chunk, so
// all 64 lanes must take part and every lane must point `x` at its query's
// inputs (lanes of one query may share a copy).  The output accumulator holds
// outputs 4g .. 4g + 3 of query n, before the output activation.
#pragma once

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "mfma_common.h"
#include "model.h"

namespace nlod {

struct NetArgs {
    const EncProgram* prog;
    const half_t* table;  // FP16 copy of the parameter vector (grid tables)
    uint32_t n_in, n_out, k_padded, n_hidden_layers;
    int32_t act, out_act;
    uint32_t packed_halves, lds_offset;  // weights' place in LDS (halves)
};

template <int W, int F, int MAXD>
__device__ __forceinline__ floatx4 run_net(const NetArgs& a, const half_t* __restrict__ sw, const float* __restrict__ x) {
    constexpr int MT = W / 16;
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const half_t* w = sw + a.lds_offset;
    const uint32_t kc0 = a.k_padded / 32;
    floatx4 acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (uint32_t c = 0; c < kc0; ++c) {
        const half8 b = encode_window<F, MAXD, half_t>(*a.prog, a.table, x, 32 * c + 8 * g);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const half8 wf = *reinterpret_cast<const half8*>(w + ((size_t)(mt * kc0 + c) * 64 + lane) * 8);
            acc[mt] = mfma16(wf, b, acc[mt]);
        }
    }
    return mlp_tail16<W>(w + (size_t)MT * kc0 * 512, acc, a.n_hidden_layers, a.act);
}

// Kernel arguments of one FP16 model (inference parameters), its packed
// weights placed at lds_offset halves.
inline NetArgs net_args_of(const Model& m, uint32_t lds_offset) {
    NetArgs a{};
    a.prog = static_cast<const EncProgram*>(m.program.p);
    a.table = static_cast<const half_t*>(m.table_half(true));
    a.n_in = m.L.n_input_dims;
    a.n_out = m.L.net.n_output;
    a.k_padded = k_padded_of(m.L.net);
    a.n_hidden_layers = m.L.net.n_hidden_layers;
    a.act = m.L.net.activation;
    a.out_act = m.L.net.output_activation;
    a.packed_halves = (uint32_t)(mfma_packed_bytes(m.L.net) / sizeof(half_t));
    a.lds_offset = lds_offset;
    return a;
}

// Calls f(W, F, MAXD) as integral constants for a pair of networks of width
// W in {32, 64} whose grids have F features per level and whose SH degree is
// at most MAXD.
template <typename Fn>
void dispatch_wf(uint32_t W, uint32_t F, uint32_t D, Fn&& f) {
    auto with_d = [&](auto Wc, auto Fc) {
        if (D <= 4) f(Wc, Fc, std::integral_constant<int, 4>{});
        else f(Wc, Fc, std::integral_constant<int, 8>{});
    };
    auto with_f = [&](auto Wc) {
        switch (F) {
            case 1: with_d(Wc, std::integral_constant<int, 1>{}); break;
            case 2: with_d(Wc, std::integral_constant<int, 2>{}); break;
            case 4: with_d(Wc, std::integral_constant<int, 4>{}); break;
            case 8: with_d(Wc, std::integral_constant<int, 8>{}); break;
            default: fail(NLOD_ERR_UNSUPPORTED, "unsupported n_features_per_level");
        }
    };
    switch (W) {
        case 32: with_f(std::integral_constant<int, 32>{}); break;
        case 64: with_f(std::integral_constant<int, 64>{}); break;
        default: fail(NLOD_ERR_UNSUPPORTED, "the fused two-network kernels support widths 32 and 64");
    }
}

// Features per grid level (2 when the layout has no grid) and the largest SH degree.
inline void layout_features(const Layout& L, uint32_t& F, uint32_t& maxdeg) {
    F = 0;
    maxdeg = 0;
    for (const Segment& s : L.segs) {
        if (s.type == SEG_GRID) F = s.F;
        if (s.type == SEG_SH) maxdeg = std::max(maxdeg, s.degree);
    }
    if (F == 0) F = 2;
}

// True when both FP16 models can share one fused launch: same width (32 or
// 64), same features per level, and both packed weight sets fit in `lds_budget`
// bytes of LDS.
inline bool fused_pair_ok(const Model& v, const Model& t, size_t lds_budget) {
    if (v.precision != PREC_FP16 || t.precision != PREC_FP16) return false;
    uint32_t Fv, Ft, dv, dt;
    layout_features(v.L, Fv, dv);
    layout_features(t.L, Ft, dt);
    const uint32_t W = v.L.net.width;
    if (!(W == t.L.net.width && (W == 32 || W == 64) && Fv == Ft)) return false;
    return mfma_packed_bytes(v.L.net) + mfma_packed_bytes(t.L.net) <= lds_budget;
}

}  // namespace nlod
