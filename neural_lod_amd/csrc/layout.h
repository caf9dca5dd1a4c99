// Model layout: the flattened encoding (segments of the encoded feature
// vector) and the MLP shape of a tcnn-style NetworkWithInputEncoding.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "json.h"

namespace nlod {

constexpr int kMaxLevels = 32;

enum SegType : int32_t { SEG_GRID = 0, SEG_SH = 1, SEG_IDENTITY = 2 };
enum GridType : int32_t { GRID_HASH = 0, GRID_DENSE = 1, GRID_TILED = 2 };
enum Interp : int32_t { INTERP_LINEAR = 0, INTERP_NEAREST = 1, INTERP_SMOOTHSTEP = 2 };
enum Activation : int32_t { ACT_NONE = 0, ACT_RELU = 1, ACT_SIGMOID = 2, ACT_EXP = 3 };

// Per-level constants of one grid, computed once on the host (so every
// kernel and the parameter offsets agree on the same float values).
struct GridLevels {
    uint32_t n_levels;
    uint32_t F;
    int32_t grid_type;
    int32_t interp;
    uint32_t offset[kMaxLevels + 1];  // entry offset of each level (F params per entry)
    uint32_t res[kMaxLevels];         // grid_resolution(scale)
    float scale[kMaxLevels];          // grid_scale(level)
};

struct Segment {
    int32_t type = SEG_GRID;
    uint32_t in_offset = 0, n_in = 0;     // input columns read
    uint32_t out_offset = 0, n_out = 0;   // first encoded row, rows of features
    uint32_t n_pad = 0;                   // padding rows (grid: 0 after; SH: 1 before; identity: 1 after)
    // grid
    uint32_t n_levels = 0, F = 0, log2_hashmap_size = 0, base_resolution = 0;
    float per_level_scale = 0.f;
    int32_t grid_type = GRID_HASH, interp = INTERP_LINEAR;
    uint64_t param_offset = 0;  // into the model's parameter vector (floats)
    uint64_t n_params = 0;
    GridLevels levels{};
    bool first_use = true;  // false for repetitions sharing an earlier segment's table
    // spherical harmonics
    uint32_t degree = 0;
    // identity
    float id_scale = 1.f, id_offset = 0.f;
};

struct NetShape {
    uint32_t width = 64, n_hidden_layers = 1, n_input = 0, n_output = 1, n_output_padded = 16;
    Activation activation = ACT_RELU, output_activation = ACT_NONE;
    uint64_t n_params() const {
        return (uint64_t)width * n_input + (uint64_t)(n_hidden_layers - 1) * width * width +
               (uint64_t)n_output_padded * width;
    }
    // parameter offset of weight matrix `layer` (0 = input, n_hidden_layers = output)
    uint64_t matrix_offset(uint32_t layer) const {
        if (layer == 0) return 0;
        return (uint64_t)width * n_input + (uint64_t)(layer - 1) * width * width;
    }
};

struct Layout {
    uint32_t n_input_dims = 0;
    uint32_t enc_width = 0;
    std::vector<Segment> segs;
    NetShape net;
    uint64_t n_net_params = 0;
    uint64_t n_params = 0;
    Json encoding_cfg, network_cfg;
    std::string describe() const;  // JSON description (tests compare it with the oracle)
};

Layout build_layout(uint32_t n_input_dims, uint32_t n_output_dims, const Json& encoding, const Json& network);
GridLevels compute_grid_levels(uint32_t n_levels, uint32_t F, uint32_t log2T, uint32_t base, float per_level_scale,
                               int32_t grid_type, int32_t interp, uint64_t* n_params_out);
Activation parse_activation(const std::string& s);

}  // namespace nlod
