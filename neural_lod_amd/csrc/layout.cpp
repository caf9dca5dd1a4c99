// Encoding-tree layout rules of tiny-cuda-nn (paths relative to
// rptr/ext/tiny-cuda-nn/):
//   create_encoding + alignment 8        src/encoding.cu:85-162
//   NetworkWithInputEncoding alignment   include/tiny-cuda-nn/network_with_input_encoding.h:44-51,
//                                        src/network.cu:111-130 (FullyFusedMLP: 16)
//   Encoding::set_alignment               include/tiny-cuda-nn/encoding.h:83-85
//   Composite nested dims / alignment    encodings/composite.h:138-215, 370-386
//   RepeatedComposite (neural_lod)       encodings/repeated_composite.h:47-135, 293-323
//   grid level table                     encodings/grid.h:958-1022, 195-204
//   grid factory defaults                encodings/grid.h:1433-1494
//   FullyFusedMLP shapes                 src/fully_fused_mlp.cu:635-678, src/network.cu:133-175
#include <cmath>
#include <memory>
#include <numeric>
#include <sstream>

#include "layout.h"

namespace nlod {
namespace {

uint32_t next_multiple(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }
uint32_t lcm_u(uint32_t a, uint32_t b) { return a / std::gcd(a, b) * b; }

struct Enc {
    virtual ~Enc() = default;
    virtual uint32_t padded_output_width() const = 0;
    virtual void set_padded_output_width(uint32_t w) = 0;
    virtual uint32_t required_output_alignment() const = 0;
    virtual uint64_t n_params() const = 0;
    // appends segments; returns the next output row
    virtual uint32_t flatten(uint32_t in_off, uint32_t out_off, uint64_t param_off, bool first_use,
                             std::vector<Segment>& out) const = 0;
    void set_alignment(uint32_t a) {
        set_padded_output_width(next_multiple(padded_output_width(), lcm_u(a, required_output_alignment())));
    }
};

std::unique_ptr<Enc> make_encoding(const Json& cfg, uint32_t n_dims, uint32_t alignment);

struct LeafEnc : Enc {
    Segment s;
    LeafEnc(const Json& cfg, uint32_t n_dims) {
        const std::string ot = lower(cfg.value_str("otype", "OneBlob"));
        s.n_in = n_dims;
        if (ot == "hashgrid" || ot == "grid" || ot == "densegrid" || ot == "tiledgrid") {
            s.type = SEG_GRID;
            if (n_dims != 3) fail(NLOD_ERR_UNSUPPORTED, "grid encodings of " + std::to_string(n_dims) + " dims are not supported (3 only)");
            const std::string dflt = ot == "tiledgrid" ? "tiled" : (ot == "densegrid" ? "dense" : "hash");
            const std::string gt = lower(cfg.value_str("type", dflt));
            s.grid_type = gt == "hash" ? GRID_HASH : gt == "dense" ? GRID_DENSE : gt == "tiled" ? GRID_TILED : -1;
            if (s.grid_type < 0) fail(NLOD_ERR_CONFIG, "invalid grid type " + gt);
            const std::string hash = lower(cfg.value_str("hash", "CoherentPrime"));
            if (hash != "coherentprime") fail(NLOD_ERR_UNSUPPORTED, "only the CoherentPrime hash is supported");
            s.F = cfg.value_u32("n_features_per_level", 2);
            if (s.F != 1 && s.F != 2 && s.F != 4 && s.F != 8)
                fail(NLOD_ERR_CONFIG, "GridEncoding: n_features_per_level must be 1, 2, 4, or 8.");
            if (cfg.contains("n_features") || cfg.contains("n_grid_features")) {
                if (cfg.contains("n_levels"))
                    fail(NLOD_ERR_CONFIG, "GridEncoding: may not specify n_features and n_levels simultaneously");
                uint32_t nf = cfg.contains("n_features") ? cfg.value_u32("n_features", 0) : cfg.value_u32("n_grid_features", 0);
                if (nf % s.F) fail(NLOD_ERR_CONFIG, "GridEncoding: n_features must be a multiple of n_features_per_level");
                s.n_levels = nf / s.F;
            } else {
                s.n_levels = cfg.value_u32("n_levels", 16);
            }
            if (s.n_levels == 0 || s.n_levels > (uint32_t)kMaxLevels)
                fail(NLOD_ERR_CONFIG, "GridEncoding: n_levels must be in [1, 32]");
            s.log2_hashmap_size = cfg.value_u32("log2_hashmap_size", 19);
            if (s.log2_hashmap_size > 30) fail(NLOD_ERR_CONFIG, "log2_hashmap_size above 30 is not supported");
            s.base_resolution = cfg.value_u32("base_resolution", 16);
            s.per_level_scale = (float)cfg.value("per_level_scale", 2.0);
            if (cfg.value_bool("stochastic_interpolation", false))
                fail(NLOD_ERR_UNSUPPORTED, "stochastic_interpolation is not supported");
            const std::string ip = lower(cfg.value_str("interpolation", "Linear"));
            s.interp = ip == "linear" ? INTERP_LINEAR : ip == "nearest" ? INTERP_NEAREST : ip == "smoothstep" ? INTERP_SMOOTHSTEP : -1;
            if (s.interp < 0) fail(NLOD_ERR_CONFIG, "invalid interpolation " + ip);
            s.levels = compute_grid_levels(s.n_levels, s.F, s.log2_hashmap_size, s.base_resolution, s.per_level_scale,
                                           s.grid_type, s.interp, &s.n_params);
            s.n_out = s.n_levels * s.F;
        } else if (ot == "sphericalharmonics") {
            s.type = SEG_SH;
            if (n_dims != 3) fail(NLOD_ERR_CONFIG, "Can only encode 3D directions in spherical harmonics.");
            s.degree = cfg.value_u32("degree", 4);
            if (s.degree == 0) fail(NLOD_ERR_CONFIG, "Spherical harmonics must have positive degree.");
            if (s.degree > 8) fail(NLOD_ERR_CONFIG, "Spherical harmonics are only implemented up to degree 8.");
            s.n_out = s.degree * s.degree;
        } else if (ot == "identity") {
            s.type = SEG_IDENTITY;
            s.n_out = n_dims;
            s.id_scale = (float)cfg.value("scale", 1.0);
            s.id_offset = (float)cfg.value("offset", 0.0);
        } else {
            fail(NLOD_ERR_UNSUPPORTED, "encoding '" + cfg.value_str("otype", "OneBlob") + "' is not supported");
        }
    }
    uint32_t padded_output_width() const override { return s.n_out + s.n_pad; }
    void set_padded_output_width(uint32_t w) override {
        if (w < s.n_out) fail(NLOD_ERR_CONFIG, "padded output width below output width");
        s.n_pad = w - s.n_out;
    }
    uint32_t required_output_alignment() const override { return s.type == SEG_GRID ? s.F : 1; }
    uint64_t n_params() const override { return s.type == SEG_GRID ? s.n_params : 0; }
    uint32_t flatten(uint32_t in_off, uint32_t out_off, uint64_t param_off, bool first_use,
                     std::vector<Segment>& out) const override {
        Segment t = s;
        t.in_offset = in_off;
        t.out_offset = out_off;
        t.param_offset = param_off;
        t.first_use = first_use;
        out.push_back(t);
        return out_off + padded_output_width();
    }
};

struct CompositeEnc : Enc {
    std::vector<std::unique_ptr<Enc>> nested;
    std::vector<uint32_t> begins;
    uint32_t reps = 1;
    CompositeEnc(const Json& cfg, uint32_t n_dims, uint32_t n_reps) : reps(n_reps) {
        if (!cfg.contains("nested") || !cfg.at("nested").is_array())
            fail(NLOD_ERR_CONFIG, "Must provide an array of nested encodings to CompositeEncoding.");
        const std::string red = lower(cfg.value_str("reduction", "Concatenation"));
        if (red != "concatenation") fail(NLOD_ERR_UNSUPPORTED, "only Concatenation reduction is supported");
        const auto& arr = cfg.at("nested").arr;
        uint64_t total = 0;
        for (auto& n : arr) {
            if (n.contains("dims_to_encode_begin")) fail(NLOD_ERR_UNSUPPORTED, "dims_to_encode_begin is not supported");
            total += n.value_u32("n_dims_to_encode", 0);
        }
        total *= reps;
        if (reps > 1 && total != n_dims)
            fail(NLOD_ERR_CONFIG, "RepeatedCompositeEncoding: nested encodings must encode exactly the number of dims of the composite");
        if (total > n_dims) fail(NLOD_ERR_CONFIG, "CompositeEncoding: nested encodings must not encode more dims than composite");
        int64_t unspecified = (int64_t)n_dims - (int64_t)total;
        uint32_t off = 0;
        for (uint32_t rep = 0; rep < reps; ++rep) {
            for (auto& n : arr) {
                uint32_t nd;
                if (n.contains("n_dims_to_encode")) {
                    nd = n.value_u32("n_dims_to_encode", 0);
                } else {
                    if (unspecified < 0)
                        fail(NLOD_ERR_CONFIG, "CompositeEncoding: may only leave 'n_dims_to_encode' unspecified for a single nested encoding");
                    nd = (uint32_t)unspecified;
                    unspecified = -1;
                }
                if (nd > 0) {
                    if (rep == 0) nested.push_back(make_encoding(n, nd, 1));
                    begins.push_back(off);
                }
                off += nd;
            }
        }
        if (nested.empty()) fail(NLOD_ERR_CONFIG, "composite encoding has no nested encoding");
        uint32_t so_far = 0;
        for (size_t i = 0; i + 1 < nested.size(); ++i) {
            uint32_t desired = nested[i + 1]->required_output_alignment();
            uint32_t need = next_multiple(so_far + nested[i]->padded_output_width(), desired) - so_far;
            nested[i]->set_padded_output_width(need);
            so_far += nested[i]->padded_output_width();
        }
    }
    uint32_t padded_output_width() const override {
        uint32_t t = 0;
        for (auto& n : nested) t += n->padded_output_width();
        return t * reps;
    }
    void set_padded_output_width(uint32_t w) override {
        uint32_t prev = padded_output_width() - reps * nested.back()->padded_output_width();
        if (w < prev || (w - prev) % reps) fail(NLOD_ERR_CONFIG, "invalid padded width for composite encoding");
        nested.back()->set_padded_output_width((w - prev) / reps);
    }
    uint32_t required_output_alignment() const override {
        uint32_t a = 1;
        for (auto& n : nested) a = lcm_u(a, n->required_output_alignment());
        return a;
    }
    uint64_t n_params() const override {
        uint64_t t = 0;
        for (auto& n : nested) t += n->n_params();
        return t;
    }
    uint32_t flatten(uint32_t in_off, uint32_t out_off, uint64_t param_off, bool first_use,
                     std::vector<Segment>& out) const override {
        size_t k = 0;
        for (uint32_t rep = 0; rep < reps; ++rep) {
            uint64_t p = param_off;
            for (auto& n : nested) {
                out_off = n->flatten(in_off + begins[k], out_off, p, first_use && rep == 0, out);
                p += n->n_params();
                ++k;
            }
        }
        return out_off;
    }
};

std::unique_ptr<Enc> make_encoding(const Json& cfg, uint32_t n_dims, uint32_t alignment) {
    if (!cfg.is_object()) fail(NLOD_ERR_CONFIG, "encoding configuration must be a JSON object");
    const std::string ot = lower(cfg.value_str("otype", "OneBlob"));
    std::unique_ptr<Enc> e;
    if (ot == "composite") e = std::make_unique<CompositeEnc>(cfg, n_dims, 1);
    else if (ot == "repeatedcomposite") e = std::make_unique<CompositeEnc>(cfg, n_dims, cfg.value_u32("n_repetitions", 1));
    else e = std::make_unique<LeafEnc>(cfg, n_dims);
    if (alignment > 0) e->set_alignment(alignment);
    return e;
}

}  // namespace

Activation parse_activation(const std::string& s) {
    const std::string a = lower(s);
    if (a == "none") return ACT_NONE;
    if (a == "relu") return ACT_RELU;
    if (a == "sigmoid") return ACT_SIGMOID;
    if (a == "exponential") return ACT_EXP;
    fail(NLOD_ERR_UNSUPPORTED, "activation '" + s + "' is not supported (None, ReLU, Sigmoid, Exponential)");
}

GridLevels compute_grid_levels(uint32_t n_levels, uint32_t F, uint32_t log2T, uint32_t base, float per_level_scale,
                               int32_t grid_type, int32_t interp, uint64_t* n_params_out) {
    GridLevels g{};
    g.n_levels = n_levels;
    g.F = F;
    g.grid_type = grid_type;
    g.interp = interp;
    const float l2s = std::log2(per_level_scale);  // float overload, as the reference's host code
    uint32_t off = 0;
    for (uint32_t l = 0; l < n_levels; ++l) {
        const float scale = std::exp2((float)l * l2s) * (float)base - 1.0f;
        const uint32_t res = (uint32_t)std::ceil(scale) + 1u;
        const uint32_t max_params = 0xffffffffu / 2u;
        uint32_t params = std::pow((float)res, 3.0f) > (float)max_params ? max_params : res * res * res;
        params = next_multiple(params, 8u);
        if (grid_type == GRID_TILED) params = std::min(params, base * base * base);
        else if (grid_type == GRID_HASH) params = std::min(params, 1u << log2T);
        g.offset[l] = off;
        g.res[l] = res;
        g.scale[l] = scale;
        off += params;
    }
    g.offset[n_levels] = off;
    if (n_params_out) *n_params_out = (uint64_t)off * F;
    return g;
}

Layout build_layout(uint32_t n_input_dims, uint32_t n_output_dims, const Json& encoding, const Json& network) {
    if (n_input_dims == 0) fail(NLOD_ERR_INVALID_ARGUMENT, "n_input_dims must be positive");
    if (n_output_dims == 0 || n_output_dims > 16) fail(NLOD_ERR_UNSUPPORTED, "n_output_dims must be in [1, 16]");
    if (!network.is_object()) fail(NLOD_ERR_CONFIG, "network configuration must be a JSON object");
    Layout L;
    L.n_input_dims = n_input_dims;
    L.encoding_cfg = encoding;
    L.network_cfg = network;
    const std::string ot = lower(network.value_str("otype", "MLP"));
    if (ot != "fullyfusedmlp" && ot != "megakernelmlp" && ot != "mlp" && ot != "cutlassmlp")
        fail(NLOD_ERR_UNSUPPORTED, "network '" + network.value_str("otype", "MLP") + "' is not supported");
    NetShape& n = L.net;
    n.width = network.value_u32("n_neurons", 128);
    if (n.width != 16 && n.width != 32 && n.width != 64 && n.width != 128)
        fail(NLOD_ERR_CONFIG, "FullyFusedMLP only supports 16, 32, 64, and 128 neurons, but got " + std::to_string(n.width));
    n.n_hidden_layers = network.value_u32("n_hidden_layers", 5);
    if (n.n_hidden_layers == 0) fail(NLOD_ERR_CONFIG, "FullyFusedMLP requires at least 1 hidden layer (3 layers in total).");
    if (n.n_hidden_layers > 16) fail(NLOD_ERR_UNSUPPORTED, "more than 16 hidden layers are not supported");
    n.activation = parse_activation(network.value_str("activation", "ReLU"));
    n.output_activation = parse_activation(network.value_str("output_activation", "None"));
    if (n.activation != ACT_RELU && n.activation != ACT_NONE)
        fail(NLOD_ERR_UNSUPPORTED, "hidden activation must be ReLU or None");

    auto enc = make_encoding(encoding, n_input_dims, 8);
    enc->set_alignment(16);  // FullyFusedMLP REQUIRED_ALIGNMENT
    L.enc_width = enc->padded_output_width();
    if (L.enc_width > 512) fail(NLOD_ERR_UNSUPPORTED, "encoded width above 512 is not supported");
    n.n_input = L.enc_width;
    n.n_output = n_output_dims;
    n.n_output_padded = next_multiple(n_output_dims, 16);
    L.n_net_params = n.n_params();
    enc->flatten(0, 0, L.n_net_params, true, L.segs);
    L.n_params = L.n_net_params + enc->n_params();
    for (auto& s : L.segs)
        if (s.in_offset + s.n_in > n_input_dims) fail(NLOD_ERR_CONFIG, "encoding reads past the input width");
    return L;
}

std::string Layout::describe() const {
    std::ostringstream o;
    o << "{\"n_input_dims\":" << n_input_dims << ",\"enc_width\":" << enc_width << ",\"n_params\":" << n_params
      << ",\"n_net_params\":" << n_net_params << ",\"width\":" << net.width << ",\"n_hidden_layers\":" << net.n_hidden_layers
      << ",\"n_output\":" << net.n_output << ",\"n_output_padded\":" << net.n_output_padded << ",\"segments\":[";
    for (size_t i = 0; i < segs.size(); ++i) {
        const Segment& s = segs[i];
        if (i) o << ",";
        o << "{\"type\":" << s.type << ",\"in_offset\":" << s.in_offset << ",\"n_in\":" << s.n_in
          << ",\"out_offset\":" << s.out_offset << ",\"n_out\":" << s.n_out << ",\"n_pad\":" << s.n_pad
          << ",\"param_offset\":" << s.param_offset << ",\"n_params\":" << (s.type == SEG_GRID ? s.n_params : 0);
        if (s.type == SEG_GRID) {
            o << ",\"levels\":[";
            for (uint32_t l = 0; l <= s.n_levels; ++l) o << (l ? "," : "") << s.levels.offset[l];
            o << "],\"res\":[";
            for (uint32_t l = 0; l < s.n_levels; ++l) o << (l ? "," : "") << s.levels.res[l];
            o << "]";
        }
        o << "}";
    }
    o << "]}";
    return o.str();
}

}  // namespace nlod
