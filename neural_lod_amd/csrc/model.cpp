// NetworkWithInputEncoding (tiny-cuda-nn include/tiny-cuda-nn/network_with_input_encoding.h)
// and its C ABI: parameters are bound by pointer (network first, then the
// encoding, network_with_input_encoding.h:113-128); inference, forward with a
// saved context, and backward producing parameter / input gradients.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>

#include "model.h"
#include "rng.h"

namespace nlod {

Model::Model(uint32_t n_in, uint32_t n_out, const Json& enc, const Json& net, int prec) : precision(prec) {
    if (prec != PREC_FP32 && prec != PREC_FP16) fail(NLOD_ERR_INVALID_ARGUMENT, "precision must be FP32 (0) or FP16 (1)");
    L = build_layout(n_in, n_out, enc, net);
    for (DevBuf* b : {&packed, &packed_bwd, &table16, &packed_inference, &packed_bwd_inference, &table16_inference,
                      &program, &ws_feat, &ws_pos, &ws_enc, &ws_delta, &ws_delta_out, &ws_denc, &ws_wgrad,
                      &ws_hidden, &ws_out})
        b->epoch = &epoch;
    if (L.n_params > L.n_net_params) {
        grid_bwd = std::make_unique<GridBackward>(L);
        grid_bwd->track(&epoch);
    }
    if (precision == PREC_FP16) {
        if (!mfma_supported(L)) fail(NLOD_ERR_UNSUPPORTED, "this layout is not supported by the FP16 MFMA kernels");
        program.ensure(mfma_program_bytes());
        mfma_upload_program(L, program.p);
    }
}

void Model::set_params(hipStream_t s, const float* p, const float* p_inf) {
    NLOD_CHECK(p != nullptr, "params pointer is null");
    ++epoch;  // graphs that read the previous binding (weights or their FP16 copies) are stale
    params = p;
    params_inference = p_inf ? p_inf : p;
    inference_aliased = params_inference == params;
    if (precision != PREC_FP16) return;
    packed.ensure(mfma_packed_bytes(L.net));
    packed_bwd.ensure(mfma_packed_bwd_bytes(L.net));
    table16.ensure(sizeof(uint16_t) * L.n_params);
    mfma_to_half(s, params, table16.p, L.n_params);
    if (!inference_aliased) {
        packed_inference.ensure(mfma_packed_bytes(L.net));
        packed_bwd_inference.ensure(mfma_packed_bwd_bytes(L.net));
        table16_inference.ensure(sizeof(uint16_t) * L.n_params);
        mfma_to_half(s, params_inference, table16_inference.p, L.n_params);
    }
    refresh_packed(s, true, !inference_aliased);
}

void Model::refresh_packed(hipStream_t s, bool training, bool inference) {
    if (precision != PREC_FP16) return;
    if (training) {
        mfma_pack_weights(s, L.net, params, packed.p, packed_bwd.p);
    }
    if (inference && !inference_aliased) {
        mfma_pack_weights(s, L.net, params_inference, packed_inference.p, packed_bwd_inference.p);
    }
}

void Model::require_params(bool inference) const {
    if (!(inference ? params_inference : params))
        fail(NLOD_ERR_STATE, "parameters are not set (nlod_model_set_params)");
}

void Model::encode(hipStream_t s, uint32_t N, const float* in, float* enc_soa, bool inference) {
    require_params(inference);
    encode_forward_f32(s, L, weights(inference), in, N, enc_soa);
}

void Model::forward_outputs(hipStream_t s, uint32_t N, const float* in, float* out, bool use_inference) {
    if (N == 0) return;
    // chunks bound the feature / encoding workspace
    const uint32_t chunk = kRecomputeAbove;
    if (precision == PREC_FP16) {
        ws_feat.ensure(mfma_feature_bytes(L, std::min(N, chunk)));
        void* pos = pos_scratch(std::min(N, chunk));
        for (uint32_t q0 = 0; q0 < N; q0 += chunk) {
            const uint32_t n = std::min(chunk, N - q0);
            mfma_encode(s, L, program.p, table_half(use_inference), in + (size_t)q0 * L.n_input_dims, n, ws_feat.p,
                        pos);
            mfma_mlp_forward(s, L, packed_fwd(use_inference), ws_feat.p, n, out + (size_t)q0 * L.net.n_output);
        }
        return;
    }
    ws_enc.ensure(sizeof(float) * (size_t)L.enc_width * std::min(N, chunk));
    for (uint32_t q0 = 0; q0 < N; q0 += chunk) {
        const uint32_t n = std::min(chunk, N - q0);
        encode_forward_f32(s, L, weights(use_inference), in + (size_t)q0 * L.n_input_dims, n, ws_enc.as<float>());
        mlp_forward_f32(s, L.net, weights(use_inference), ws_enc.as<float>(), n, out + (size_t)q0 * L.net.n_output,
                        nullptr);
    }
}

void* Model::pos_scratch(uint32_t N) {
    const size_t bytes = mfma_pos_scratch_bytes(L, N);
    if (bytes == 0) return nullptr;
    ws_pos.ensure(bytes);
    return ws_pos.p;
}

void Model::inference(hipStream_t s, uint32_t N, const float* in, float* out) {
    require_params(true);
    forward_outputs(s, N, in, out, true);
}

static size_t enc_bytes(const Layout& L, int precision, uint32_t N) {
    return precision == PREC_FP16 ? mfma_feature_bytes(L, N) : sizeof(float) * (size_t)L.enc_width * N;
}
static size_t hidden_bytes(const Layout& L, int precision, uint32_t N) {
    return precision == PREC_FP16 ? mfma_hidden_bytes(L, N)
                                  : sizeof(float) * (size_t)L.net.n_hidden_layers * L.net.width * N;
}

void Model::forward_saved(hipStream_t s, uint32_t N, const float* in, bool use_inference, void* enc, void* hidden,
                          float* out) {
    if (precision == PREC_FP16) {
        mfma_encode(s, L, program.p, table_half(use_inference), in, N, enc, pos_scratch(N));
        mfma_mlp_forward_train(s, L, packed_fwd(use_inference), enc, N, hidden, out);
        return;
    }
    encode_forward_f32(s, L, weights(use_inference), in, N, static_cast<float*>(enc));
    mlp_forward_f32(s, L.net, weights(use_inference), static_cast<const float*>(enc), N, out,
                    static_cast<float*>(hidden));
}

std::unique_ptr<ForwardCtx> Model::forward(hipStream_t s, uint32_t N, const float* in, float* out, bool use_inference,
                                           bool prepare_input_gradients) {
    auto ctx = std::make_unique<ForwardCtx>();
    forward_into(s, N, in, out, use_inference, prepare_input_gradients, *ctx);
    return ctx;
}

void Model::forward_into(hipStream_t s, uint32_t N, const float* in, float* out, bool use_inference,
                         bool prepare_input_gradients, ForwardCtx& ctx) {
    require_params(use_inference);
    ctx.N = N;
    ctx.use_inference_params = use_inference;
    ctx.prepare_input_gradients = prepare_input_gradients;
    ctx.recompute = N > kRecomputeAbove;
    if (N == 0) return;
    const size_t out_bytes = sizeof(float) * (size_t)L.net.n_output * N;
    ctx.out.ensure(out_bytes);
    if (ctx.recompute) {
        // activations are recomputed chunk by chunk in backward
        forward_outputs(s, N, in, ctx.out.as<float>(), use_inference);
    } else {
        ctx.enc.ensure(enc_bytes(L, precision, N));
        ctx.hidden.ensure(hidden_bytes(L, precision, N));
        forward_saved(s, N, in, use_inference, ctx.enc.p, ctx.hidden.p, ctx.out.as<float>());
    }
    if (out) NLOD_HIP(hipMemcpyAsync(out, ctx.out.p, out_bytes, hipMemcpyDeviceToDevice, s));
}

void Model::grid_backward(hipStream_t s, const float* in, uint32_t N, const float* d_enc, float* d_params,
                          bool overwrite) {
    if (grid_overwrite_covers_tables()) {
        grid_bwd->run(s, in, N, d_enc, d_params, overwrite);
        return;
    }
    if (overwrite) fill_f32(s, d_params + L.n_net_params, L.n_params - L.n_net_params, 0.0f);
    encode_backward_params(s, L, in, N, d_enc, d_params);
}

void Model::backward_saved(hipStream_t s, uint32_t N, bool use_inference, const float* in, const float* d_out,
                           const void* enc, const void* hidden, const float* out, float* d_params, float* d_in,
                           bool param_grads, bool net_accumulate, bool grid_overwrite) {
    const NetShape& n = L.net;
    const bool grid_grads = param_grads && L.n_params > L.n_net_params;
    const bool need_denc = grid_grads || d_in;
    if (need_denc) ws_denc.ensure(sizeof(float) * (size_t)L.enc_width * N);
    float* d_enc = need_denc ? ws_denc.as<float>() : nullptr;
    if (precision == PREC_FP16) {
        ws_delta.ensure(mfma_delta_bytes(L, N));
        mfma_mlp_backward(s, L, packed_back(use_inference), hidden, out, d_out, N, ws_delta.p, d_enc);
        if (param_grads) {
            const size_t wsf = mfma_wgrad_workspace_floats(L, N);
            ws_wgrad.ensure(sizeof(float) * wsf);
            mfma_weight_grad(s, L, enc, hidden, ws_delta.p, N, d_params, ws_wgrad.as<float>(), wsf, net_accumulate);
        }
    } else {
        const float* w = weights(use_inference);
        const float* h = static_cast<const float*>(hidden);
        ws_delta.ensure(sizeof(float) * (size_t)n.n_hidden_layers * n.width * N);
        ws_delta_out.ensure(sizeof(float) * (size_t)n.n_output_padded * N);
        mlp_backward_delta_f32(s, n, w, h, out, d_out, N, ws_delta.as<float>(), ws_delta_out.as<float>(), d_enc);
        if (param_grads) {
            const size_t wsf = mlp_weight_grad_workspace_floats(n, N);
            ws_wgrad.ensure(sizeof(float) * wsf);
            mlp_weight_grad_f32(s, n, static_cast<const float*>(enc), h, ws_delta.as<float>(),
                                ws_delta_out.as<float>(), N, d_params, ws_wgrad.as<float>(), wsf, net_accumulate);
        }
    }
    if (grid_grads) grid_backward(s, in, N, d_enc, d_params, grid_overwrite);
    if (d_in) encode_backward_input(s, L, in, N, d_enc, d_in);
}

void Model::backward(hipStream_t s, ForwardCtx& ctx, const float* in, const float* d_out, float* d_params, float* d_in,
                     int grad_mode, bool grid_grads_zero) {
    NLOD_CHECK(grad_mode >= GRAD_IGNORE && grad_mode <= GRAD_ACCUMULATE, "invalid gradient mode");
    NLOD_CHECK(grad_mode == GRAD_IGNORE || d_params, "parameter-gradient pointer is null");
    if (d_in && !encode_backward_input_supported(L))
        fail(NLOD_ERR_UNSUPPORTED, "input gradients through a grid encoding are not supported");
    const uint32_t N = ctx.N;
    const bool param_grads = grad_mode != GRAD_IGNORE;
    const bool has_grid = L.n_params > L.n_net_params;
    // OVERWRITE: the first batch (chunk) stores the grid gradients, later ones add
    const bool grid_overwrite = grad_mode == GRAD_OVERWRITE && !(grid_grads_zero && !grid_overwrite_covers_tables());
    if (N == 0) {
        if (grad_mode == GRAD_OVERWRITE) fill_f32(s, d_params, grid_grads_zero ? L.n_net_params : L.n_params, 0.0f);
        return;
    }
    NLOD_CHECK(in && d_out, "input / output-gradient pointer is null");
    const bool use_inf = ctx.use_inference_params;
    (void)has_grid;
    if (!ctx.recompute) {
        backward_saved(s, N, use_inf, in, d_out, ctx.enc.p, ctx.hidden.p, ctx.out.as<float>(), d_params, d_in,
                       param_grads, grad_mode == GRAD_ACCUMULATE, grid_overwrite);
        return;
    }
    // recompute the activations of each chunk, then its backward; gradients of
    // later chunks are added to the first chunk's
    const uint32_t chunk = kRecomputeAbove;
    ws_feat.ensure(enc_bytes(L, precision, chunk));
    ws_hidden.ensure(hidden_bytes(L, precision, chunk));
    ws_out.ensure(sizeof(float) * (size_t)L.net.n_output * chunk);
    for (uint32_t q0 = 0; q0 < N; q0 += chunk) {
        const uint32_t n = std::min(chunk, N - q0);
        const float* in_c = in + (size_t)q0 * L.n_input_dims;
        forward_saved(s, n, in_c, use_inf, ws_feat.p, ws_hidden.p, ws_out.as<float>());
        backward_saved(s, n, use_inf, in_c, d_out + (size_t)q0 * L.net.n_output, ws_feat.p, ws_hidden.p,
                       ws_out.as<float>(), d_params, d_in ? d_in + (size_t)q0 * L.n_input_dims : nullptr, param_grads,
                       grad_mode == GRAD_ACCUMULATE || q0 > 0, grid_overwrite && q0 == 0);
    }
}

uint32_t Model::train_fused(hipStream_t s, uint32_t N, const float* in, const float* target, int loss_type,
                           const float* loss_scale, double* loss_partial, uint32_t max_partials, float* d_params) {
    if (precision != PREC_FP16 || N == 0 || N > kRecomputeAbove || !mfma_fused_train_supported(L)) return 0;
    const uint32_t blocks = mfma_fused_train_blocks(L, N);
    if (blocks > max_partials) return 0;
    require_params(false);
    const bool has_grid = L.n_params > L.n_net_params;
    ws_feat.ensure(mfma_feature_bytes(L, N));
    if (has_grid) ws_denc.ensure(sizeof(float) * (size_t)L.enc_width * N);
    ws_wgrad.ensure(sizeof(float) * (size_t)blocks * L.n_net_params);
    mfma_encode(s, L, program.p, table_half(false), in, N, ws_feat.p, pos_scratch(N));
    mfma_fused_train(s, L, packed_fwd(false), packed_back(false), ws_feat.p, target, N, loss_type, loss_scale,
                     loss_partial, has_grid ? ws_denc.as<float>() : nullptr, ws_wgrad.as<float>());
    partial_reduce(s, (uint32_t)L.n_net_params, blocks, ws_wgrad.as<float>(), d_params, false);
    if (has_grid) grid_backward(s, in, N, ws_denc.as<float>(), d_params, true);
    return blocks;
}

// Parameter initialisation of tiny-cuda-nn's Trainer (trainer.h:53-92):
// pcg32 seeded from std::seed_seq{seed}; network matrices Xavier-uniform on the
// host in order (gpu_matrix.h:291-306), then each grid table U[-1e-4, 1e-4]
// drawn as generate_random_uniform does it (random.h:63-85): its launch has
// T = ceil(ceil(n/4) / 128) * 128 threads (common.h:289-293), and thread t takes
// stream values 4t..4t+3 for elements t + j*T, j = 0..3; the stream then
// advances by n.
void Model::init_params_host(uint32_t seed, float* out) const {
    std::seed_seq seq{seed};
    std::vector<uint32_t> seeds(2);
    seq.generate(seeds.begin(), seeds.end());
    Pcg32 rng(seeds.front());
    const NetShape& n = L.net;
    float* p = out;
    auto xavier = [&](uint32_t rows, uint32_t cols) {
        const float scale = std::sqrt(6.0f / (float)(cols + rows));
        for (size_t k = 0; k < (size_t)rows * cols; ++k) p[k] = rng.next_float() * 2.0f * scale - scale;
        p += (size_t)rows * cols;
    };
    xavier(n.width, n.n_input);
    for (uint32_t l = 1; l < n.n_hidden_layers; ++l) xavier(n.width, n.width);
    xavier(n.n_output_padded, n.width);
    for (const Segment& s : L.segs) {
        if (s.type != SEG_GRID || !s.first_use) continue;
        const size_t count = s.n_params;
        const size_t n_threads = ((count + 3) / 4 + 127) / 128 * 128;
        const float lo = -1e-4f, hi = 1e-4f, span = hi - lo;
        Pcg32 r = rng;
        float* t = out + s.param_offset;
        for (size_t pos = 0; pos < n_threads * 4; ++pos) {
            const float v = r.next_float();
            const size_t idx = pos / 4 + n_threads * (pos % 4);
            if (idx < count) t[idx] = std::fma(v, span, lo);
        }
        rng.advance(count);
    }
}

}  // namespace nlod

// ============================================================== C ABI
struct nlod_model {
    std::unique_ptr<nlod::Model> m;
};
struct nlod_ctx {
    std::unique_ptr<nlod::ForwardCtx> c;
    nlod_model* owner = nullptr;
};

namespace nlod {
Model& model_of(nlod_model* m) {
    NLOD_CHECK(m && m->m, "null model");
    return *m->m;
}
}  // namespace nlod

extern "C" {

int nlod_model_create(uint32_t n_input_dims, uint32_t n_output_dims, const char* encoding_json,
                      const char* network_json, int precision, nlod_model** out) {
    return nlod::guarded([&] {
        NLOD_CHECK(out && encoding_json && network_json, "null argument");
        *out = nullptr;
        const nlod::Json enc = nlod::Json::parse(encoding_json);
        const nlod::Json net = nlod::Json::parse(network_json);
        auto h = std::make_unique<nlod_model>();
        h->m = std::make_unique<nlod::Model>(n_input_dims, n_output_dims, enc, net, precision);
        *out = h.release();
    });
}

int nlod_model_destroy(nlod_model* model) {
    return nlod::guarded([&] { delete model; });
}

int nlod_model_get_info(const nlod_model* model, nlod_model_info* info) {
    return nlod::guarded([&] {
        NLOD_CHECK(model && info, "null argument");
        const nlod::Layout& L = model->m->L;
        std::memset(info, 0, sizeof(*info));
        info->n_params = L.n_params;
        info->n_network_params = L.n_net_params;
        info->input_width = L.n_input_dims;
        info->encoded_width = L.enc_width;
        info->padded_output_width = L.net.n_output_padded;
        info->output_width = L.net.n_output;
        info->network_width = L.net.width;
        info->n_hidden_layers = L.net.n_hidden_layers;
        info->precision = (uint32_t)model->m->precision;
    });
}

int nlod_model_describe(const nlod_model* model, char* buf, size_t buf_len, size_t* needed) {
    return nlod::guarded([&] {
        NLOD_CHECK(model, "null model");
        const std::string d = model->m->L.describe();
        if (needed) *needed = d.size() + 1;
        if (buf && buf_len) {
            const size_t n = std::min(buf_len - 1, d.size());
            std::memcpy(buf, d.data(), n);
            buf[n] = '\0';
        }
    });
}

int nlod_model_init_params(const nlod_model* model, uint32_t seed, float* params_host) {
    return nlod::guarded([&] {
        NLOD_CHECK(model && params_host, "null argument");
        model->m->init_params_host(seed, params_host);
    });
}

int nlod_model_set_params(nlod_model* model, void* stream, const float* d_params, const float* d_inference_params) {
    return nlod::guarded([&] {
        NLOD_CHECK(model, "null model");
        model->m->set_params(nlod::as_stream(stream), d_params, d_inference_params);
    });
}

int nlod_model_encode(nlod_model* model, void* stream, uint32_t n, const float* d_in, float* d_encoded,
                      int use_inference_params) {
    return nlod::guarded([&] {
        NLOD_CHECK(model && (n == 0 || (d_in && d_encoded)), "null argument");
        model->m->encode(nlod::as_stream(stream), n, d_in, d_encoded, use_inference_params != 0);
    });
}

int nlod_model_encode_backward(nlod_model* model, void* stream, uint32_t n, const float* d_in,
                               const float* d_dL_dencoded, float* d_dL_dparams, int grad_mode) {
    return nlod::guarded([&] {
        NLOD_CHECK(model && d_dL_dparams, "null argument");
        nlod::Model& m = *model->m;
        NLOD_CHECK(grad_mode == nlod::GRAD_OVERWRITE || grad_mode == nlod::GRAD_ACCUMULATE, "invalid gradient mode");
        hipStream_t s = nlod::as_stream(stream);
        const bool overwrite = grad_mode == nlod::GRAD_OVERWRITE;
        if (overwrite) nlod::fill_f32(s, d_dL_dparams, m.L.n_net_params, 0.0f);
        if (m.L.n_params == m.L.n_net_params) return;
        if (n) m.grid_backward(s, d_in, n, d_dL_dencoded, d_dL_dparams, overwrite);
        else if (overwrite) nlod::fill_f32(s, d_dL_dparams + m.L.n_net_params, m.L.n_params - m.L.n_net_params, 0.0f);
    });
}

int nlod_model_feature_bytes(const nlod_model* model, uint32_t n, uint64_t* bytes_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(model && bytes_out, "null argument");
        if (model->m->precision != nlod::PREC_FP16) nlod::fail(NLOD_ERR_UNSUPPORTED, "FP16 models only");
        *bytes_out = nlod::mfma_feature_bytes(model->m->L, n);
    });
}

int nlod_model_encode_fragments(nlod_model* model, void* stream, uint32_t n, const float* d_in, void* d_features) {
    return nlod::guarded([&] {
        NLOD_CHECK(model && (n == 0 || (d_in && d_features)), "null argument");
        nlod::Model& m = *model->m;
        if (m.precision != nlod::PREC_FP16) nlod::fail(NLOD_ERR_UNSUPPORTED, "FP16 models only");
        m.require_params(true);
        nlod::mfma_encode(nlod::as_stream(stream), m.L, m.program.p, m.table_half(true), d_in, n, d_features,
                          m.pos_scratch(n));
    });
}

int nlod_model_mlp_fragments(nlod_model* model, void* stream, uint32_t n, const void* d_features, float* d_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(model && (n == 0 || (d_features && d_out)), "null argument");
        nlod::Model& m = *model->m;
        if (m.precision != nlod::PREC_FP16) nlod::fail(NLOD_ERR_UNSUPPORTED, "FP16 models only");
        m.require_params(true);
        if (n) nlod::mfma_mlp_forward(nlod::as_stream(stream), m.L, m.packed_fwd(true), d_features, n, d_out);
    });
}

int nlod_model_inference(nlod_model* model, void* stream, uint32_t n, const float* d_in, float* d_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(model && (n == 0 || (d_in && d_out)), "null argument");
        model->m->inference(nlod::as_stream(stream), n, d_in, d_out);
    });
}

int nlod_model_forward(nlod_model* model, void* stream, uint32_t n, const float* d_in, float* d_out,
                       int use_inference_params, int prepare_input_gradients, nlod_ctx** ctx_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(model && ctx_out && (n == 0 || d_in), "null argument");
        *ctx_out = nullptr;
        auto c = std::make_unique<nlod_ctx>();
        c->c = model->m->forward(nlod::as_stream(stream), n, d_in, d_out, use_inference_params != 0,
                                 prepare_input_gradients != 0);
        c->owner = model;
        *ctx_out = c.release();
    });
}

int nlod_model_backward(nlod_model* model, void* stream, nlod_ctx* ctx, const float* d_in, const float* d_dL_dout,
                        float* d_dL_dparams, float* d_dL_dinput, int grad_mode) {
    return nlod::guarded([&] {
        NLOD_CHECK(model && ctx && ctx->c, "null argument");
        if (ctx->owner != model) nlod::fail(NLOD_ERR_STATE, "context belongs to another model");
        model->m->backward(nlod::as_stream(stream), *ctx->c, d_in, d_dL_dout, d_dL_dparams, d_dL_dinput, grad_mode);
    });
}

int nlod_ctx_destroy(nlod_ctx* ctx) {
    return nlod::guarded([&] { delete ctx; });
}

}  // extern "C"
