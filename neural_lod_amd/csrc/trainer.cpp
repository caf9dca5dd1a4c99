// Trainer (tiny-cuda-nn include/tiny-cuda-nn/trainer.h:50-330) on gfx950:
// FP32 master parameters, Adam moments and per-parameter step counts; one
// training step = forward, loss, backward, FP64 gradient check, optimizer,
// with dynamic loss scaling decided on the device.  Steps can be captured
// into a hipGraph per batch size (trainer.h:171-176 captures forward +
// backward in a CUDA graph; here the optimizer joins the graph because no
// decision needs the host).
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "json.h"
#include "model.h"
#include "train.h"

namespace nlod {
namespace {

struct OptimizerConfig {
    AdamParams adam{};
    float learning_rate = 1e-3f;
    bool exp_decay = false;
    float decay_base = 0.1f;
    uint32_t decay_start = 10000, decay_interval = 10000, decay_end = 10000000;
    bool ema = false;
    float ema_decay = 0.99f;
};

// create_optimizer (tiny-cuda-nn src/optimizer.cu:47-80) for the otypes the
// path uses: Adam, ExponentialDecay{nested} and Ema{nested}, nested in any order.
void parse_optimizer(const Json& j, OptimizerConfig& c, int depth = 0) {
    if (depth > 4) fail(NLOD_ERR_CONFIG, "optimizer nesting too deep");
    const std::string t = lower(j.value_str("otype", "Adam"));
    if (t == "adam") {
        AdamParams& a = c.adam;
        a.beta1 = (float)j.value("beta1", 0.9);
        a.beta2 = (float)j.value("beta2", 0.999);
        a.epsilon = (float)j.value("epsilon", 1e-8);
        a.l2_reg = (float)j.value("l2_reg", 1e-8);
        a.relative_decay = (float)j.value("relative_decay", 0.0);
        a.absolute_decay = (float)j.value("absolute_decay", 0.0);
        a.clipping = (float)j.value("clipping_magnitude", 0.0);
        a.non_matrix_lr_factor = (float)j.value("non_matrix_learning_rate_factor", 1.0);
        a.optimize_matrix = j.value_bool("optimize_matrix_params", true);
        a.optimize_non_matrix = j.value_bool("optimize_non_matrix_params", true);
        a.adabound = j.value_bool("adabound", false);
        c.learning_rate = (float)j.value("learning_rate", 1e-3);
    } else if (t == "exponentialdecay") {
        if (c.exp_decay) fail(NLOD_ERR_CONFIG, "more than one ExponentialDecay optimizer");
        c.exp_decay = true;
        c.decay_base = (float)j.value("decay_base", 0.1);
        c.decay_interval = j.value_u32("decay_interval", 10000);
        c.decay_start = j.value_u32("decay_start", 10000);
        c.decay_end = j.value_u32("decay_end", 10000000);
        if (c.decay_interval == 0) fail(NLOD_ERR_CONFIG, "decay_interval must be positive");
        parse_optimizer(j.contains("nested") ? j.at("nested") : Json::parse("{}"), c, depth + 1);
    } else if (t == "ema") {
        if (c.ema) fail(NLOD_ERR_CONFIG, "more than one Ema optimizer");
        c.ema = true;
        c.ema_decay = (float)j.value("decay", 0.99);
        parse_optimizer(j.contains("nested") ? j.at("nested") : Json::parse("{}"), c, depth + 1);
    } else {
        fail(NLOD_ERR_UNSUPPORTED, "optimizer otype '" + j.value_str("otype", "") + "' is not supported");
    }
}

int parse_loss(const Json& j) {
    const std::string t = lower(j.value_str("otype", "RelativeL2"));
    if (t == "l2") return LOSS_L2;
    if (t == "relativel2") return LOSS_RELATIVE_L2;
    if (t == "binarycrossentropy") return LOSS_BCE;
    fail(NLOD_ERR_UNSUPPORTED, "loss otype '" + j.value_str("otype", "") + "' is not supported");
}

template <typename T>
struct DevArray {
    DevBuf buf;
    size_t n = 0;
    void alloc(size_t count) {
        buf.ensure(sizeof(T) * (count ? count : 1));
        n = count;
    }
    T* p() const { return buf.as<T>(); }
};

constexpr char kMagic[8] = {'N', 'L', 'O', 'D', 'T', 'R', 'N', '1'};

}  // namespace

struct Trainer {
    Model& model;
    OptimizerConfig opt;
    int loss_type = LOSS_RELATIVE_L2;
    TrainHyper hyper{};
    float initial_scale = 1024.0f;
    bool use_graphs = false;
    bool fused = true;  // FP16: fused forward + loss + backward kernel where the layout allows
    uint64_t n = 0, n_matrix = 0;

    DevArray<float> params, grads, m1, m2, ema;
    DevArray<uint32_t> steps;
    DevArray<TrainState> state;
    DevArray<double> loss_partial, grad_partial;
    DevArray<uint32_t> bad_partial;
    DevBuf d_out, stage_in, stage_target;
    ForwardCtx ctx;

    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    std::map<uint32_t, hipGraphExec_t> graphs;
    std::map<uint32_t, int> warm;  // batch sizes that ran once eagerly (workspaces sized)
    // A captured step holds the addresses of the trainer's step buffers and of
    // the model's buffers; every graph in `graphs` was captured while both
    // epochs had the values in graph_stamp.
    uint64_t own_epoch = 0;
    std::pair<uint64_t, uint64_t> graph_stamp{0, 0};
    std::pair<uint64_t, uint64_t> stamp() const { return {model.epoch, own_epoch}; }

    Trainer(Model& m, const Json& optimizer, const Json& loss, const Json& scaler, uint32_t seed, bool graphs_on,
            bool fused_on)
        : model(m), use_graphs(graphs_on), fused(fused_on) {
        parse_optimizer(optimizer, opt);
        loss_type = parse_loss(loss);
        initial_scale = (float)scaler.value("initial_scale", 1024.0);
        hyper.dynamic_scale = scaler.value_bool("dynamic", true);
        hyper.growth = (float)scaler.value("growth_factor", 2.0);
        hyper.backoff = (float)scaler.value("backoff_factor", 0.5);
        hyper.growth_interval = scaler.value_u32("growth_interval", 2000);
        hyper.min_scale = (float)scaler.value("min_scale", 1.0);
        hyper.max_scale = (float)scaler.value("max_scale", 65536.0 * 65536.0);
        if (!(initial_scale > 0.0f) || !(hyper.growth >= 1.0f) || !(hyper.backoff > 0.0f && hyper.backoff <= 1.0f) ||
            hyper.growth_interval == 0)
            fail(NLOD_ERR_CONFIG, "invalid loss-scale configuration");
        hyper.base_lr = opt.learning_rate;
        hyper.decay_base = opt.decay_base;
        hyper.decay_start = opt.decay_start;
        hyper.decay_interval = opt.exp_decay ? opt.decay_interval : 0;
        hyper.decay_end = opt.decay_end;

        n = model.L.n_params;
        n_matrix = model.L.n_net_params;
        params.alloc(n);
        grads.alloc(n);
        m1.alloc(n);
        m2.alloc(n);
        steps.alloc(n);
        if (opt.ema) ema.alloc(n);
        state.alloc(1);
        loss_partial.alloc(kLossPartials);
        grad_partial.alloc(kGradPartials);
        bad_partial.alloc(kGradPartials);
        for (DevBuf* b : {&d_out, &stage_in, &stage_target, &ctx.enc, &ctx.hidden, &ctx.out}) b->epoch = &own_epoch;
        NLOD_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        NLOD_HIP(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
        NLOD_HIP(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));

        // Trainer::initialize_params (trainer.h:71-92): model init from seed
        std::vector<float> host(n);
        model.init_params_host(seed, host.data());
        NLOD_HIP(hipMemcpyAsync(params.p(), host.data(), sizeof(float) * n, hipMemcpyHostToDevice, stream));
        reset_optimizer(stream);
        if (opt.ema) fill_f32(stream, ema.p(), n, 0.0f);  // EmaOptimizer::allocate memsets the EMA to 0
        bind(stream);
        NLOD_HIP(hipStreamSynchronize(stream));
    }

    ~Trainer() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }

    void reset_optimizer(hipStream_t s) {
        fill_f32(s, grads.p(), n, 0.0f);
        fill_f32(s, m1.p(), n, 0.0f);
        fill_f32(s, m2.p(), n, 0.0f);
        fill_f32(s, reinterpret_cast<float*>(steps.p()), n, 0.0f);  // all-zero bits
        train_state_init(s, state.p(), initial_scale);
    }

    // Trainer::reset_param_pointers (trainer.h:314-328): inference uses the EMA weights if any.
    void bind(hipStream_t s) { model.set_params(s, params.p(), opt.ema ? ema.p() : nullptr); }

    void drop_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }

    void step_body(hipStream_t s, uint32_t N, const float* in, const float* target) {
        const bool fp16 = model.precision == PREC_FP16;
        // FP16 form: one fused pass writes every gradient and the loss partials
        uint32_t n_loss = fused ? model.train_fused(s, N, in, target, loss_type, &state.p()->loss_scale,
                                                    loss_partial.p(), kLossPartials, grads.p())
                                : 0;
        if (n_loss == 0) {
            model.forward_into(s, N, in, nullptr, false, false, ctx);
            const uint64_t n_out = (uint64_t)N * model.L.net.n_output;
            loss_forward(s, loss_type, n_out, model.L.net.n_output, state.p(), ctx.out.as<float>(), target,
                         d_out.as<float>(), loss_partial.p());
            n_loss = kLossPartials;
            // the grid part of the gradient buffer is all zero here unless the grid
            // backward overwrites it (cleared by the previous optimizer step)
            model.backward(s, ctx, in, d_out.as<float>(), grads.p(), nullptr, GRAD_OVERWRITE, true);
        }
        grad_stats(s, n, grads.p(), grad_partial.p(), bad_partial.p());
        train_finalize(s, hyper, n_loss, loss_partial.p(), grad_partial.p(), bad_partial.p(), state.p());
        // the next backward overwrites every gradient when the binned grid backward
        // covers the tables; otherwise the grid part is accumulated into zeros
        adam_update(s, opt.adam, n, n_matrix, state.p(), params.p(), fp16 ? model.table16.p : nullptr, grads.p(),
                    m1.p(), m2.p(), steps.p(), !model.grid_overwrite_covers_tables());
        if (opt.ema)
            ema_update(s, opt.ema_decay, n, state.p(), params.p(), ema.p(), fp16 ? model.table16_inference.p : nullptr);
        model.refresh_packed(s, true, opt.ema);
    }

    void step(hipStream_t caller, uint32_t N, const float* in, const float* target) {
        NLOD_CHECK(N > 0 && in && target, "empty batch or null pointer");
        d_out.ensure(sizeof(float) * (size_t)N * model.L.net.n_output);
        NLOD_HIP(hipEventRecord(ev_in, caller));
        NLOD_HIP(hipStreamWaitEvent(stream, ev_in, 0));
        if (!use_graphs) {
            step_body(stream, N, in, target);
        } else {
            const size_t in_bytes = sizeof(float) * (size_t)N * model.L.n_input_dims;
            const size_t tg_bytes = sizeof(float) * (size_t)N * model.L.net.n_output;
            stage_in.ensure(in_bytes);
            stage_target.ensure(tg_bytes);
            NLOD_HIP(hipMemcpyAsync(stage_in.p, in, in_bytes, hipMemcpyDeviceToDevice, stream));
            NLOD_HIP(hipMemcpyAsync(stage_target.p, target, tg_bytes, hipMemcpyDeviceToDevice, stream));
            const float* sin = stage_in.as<float>();
            const float* stg = stage_target.as<float>();
            // a buffer that a graph captured was freed or moved since (a larger
            // inference or backward call on the model, new parameters, a larger
            // step): every graph is stale and each size runs eagerly once more
            if (!graphs.empty() && stamp() != graph_stamp) {
                drop_graphs();
                warm.clear();
            }
            auto it = graphs.find(N);
            if (it != graphs.end()) {
                NLOD_HIP(hipGraphLaunch(it->second, stream));
            } else if (!warm.count(N)) {
                // first step at this size: eager, so every workspace gets allocated outside capture
                step_body(stream, N, sin, stg);
                warm[N] = 1;
            } else {
                const auto before = stamp();
                hipGraph_t g = nullptr;
                NLOD_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
                try {
                    step_body(stream, N, sin, stg);
                } catch (...) {
                    (void)hipStreamEndCapture(stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                NLOD_HIP(hipStreamEndCapture(stream, &g));
                hipGraphExec_t exec = nullptr;
                const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                NLOD_HIP(e);
                if (stamp() != before) {
                    // a buffer moved while the body was recorded: the graph may hold a
                    // freed address, so it is dropped and this step runs eagerly
                    (void)hipGraphExecDestroy(exec);
                    step_body(stream, N, sin, stg);
                } else {
                    if (graphs.empty()) graph_stamp = before;
                    graphs[N] = exec;
                    NLOD_HIP(hipGraphLaunch(exec, stream));
                }
            }
        }
        NLOD_HIP(hipEventRecord(ev_out, stream));
        NLOD_HIP(hipStreamWaitEvent(caller, ev_out, 0));
    }

    TrainState read_state() {
        TrainState h{};
        NLOD_HIP(hipMemcpyAsync(&h, state.p(), sizeof(h), hipMemcpyDeviceToHost, stream));
        NLOD_HIP(hipStreamSynchronize(stream));
        return h;
    }

    // Trainer::set_params_full_precision (trainer.h:231-245): master, and the
    // inference (EMA) weights, from one FP32 vector.
    void set_params(hipStream_t caller, const float* src) {
        NLOD_HIP(hipEventRecord(ev_in, caller));
        NLOD_HIP(hipStreamWaitEvent(stream, ev_in, 0));
        NLOD_HIP(hipMemcpyAsync(params.p(), src, sizeof(float) * n, hipMemcpyDeviceToDevice, stream));
        if (opt.ema) NLOD_HIP(hipMemcpyAsync(ema.p(), src, sizeof(float) * n, hipMemcpyDeviceToDevice, stream));
        bind(stream);
        NLOD_HIP(hipEventRecord(ev_out, stream));
        NLOD_HIP(hipStreamWaitEvent(caller, ev_out, 0));
    }

    // Snapshot: header, inference parameters (what trainer.h:267-278 stores),
    // then optionally master parameters, moments, step counts and the state.
    size_t snapshot_bytes(bool with_opt) const {
        size_t b = sizeof(kMagic) + 2 * sizeof(uint64_t) + sizeof(float) * n;
        if (with_opt) b += sizeof(float) * 3 * n + sizeof(uint32_t) * n + sizeof(TrainState);
        return b;
    }

    void serialize(bool with_opt, uint8_t* out) {
        NLOD_HIP(hipStreamSynchronize(stream));
        uint8_t* p = out;
        std::memcpy(p, kMagic, sizeof(kMagic));
        p += sizeof(kMagic);
        const uint64_t hdr[2] = {n, with_opt ? 1ull : 0ull};
        std::memcpy(p, hdr, sizeof(hdr));
        p += sizeof(hdr);
        auto d2h = [&](const void* src, size_t bytes) {
            NLOD_HIP(hipMemcpy(p, src, bytes, hipMemcpyDeviceToHost));
            p += bytes;
        };
        d2h(opt.ema ? ema.p() : params.p(), sizeof(float) * n);
        if (with_opt) {
            d2h(params.p(), sizeof(float) * n);
            d2h(m1.p(), sizeof(float) * n);
            d2h(m2.p(), sizeof(float) * n);
            d2h(steps.p(), sizeof(uint32_t) * n);
            d2h(state.p(), sizeof(TrainState));
        }
    }

    void deserialize(const uint8_t* in, size_t len) {
        if (len < sizeof(kMagic) + 2 * sizeof(uint64_t) || std::memcmp(in, kMagic, sizeof(kMagic)) != 0)
            fail(NLOD_ERR_INVALID_ARGUMENT, "not a trainer snapshot");
        uint64_t hdr[2];
        std::memcpy(hdr, in + sizeof(kMagic), sizeof(hdr));
        if (hdr[0] != n) fail(NLOD_ERR_INVALID_ARGUMENT, "snapshot has a different parameter count");
        const bool with_opt = hdr[1] != 0;
        if (len != snapshot_bytes(with_opt)) fail(NLOD_ERR_INVALID_ARGUMENT, "snapshot has the wrong size");
        NLOD_HIP(hipStreamSynchronize(stream));
        const uint8_t* p = in + sizeof(kMagic) + sizeof(hdr);
        auto h2d = [&](void* dst, size_t bytes) {
            NLOD_HIP(hipMemcpy(dst, p, bytes, hipMemcpyHostToDevice));
            p += bytes;
        };
        const uint8_t* inference = p;
        h2d(params.p(), sizeof(float) * n);
        if (opt.ema) NLOD_HIP(hipMemcpy(ema.p(), inference, sizeof(float) * n, hipMemcpyHostToDevice));
        if (with_opt) {
            h2d(params.p(), sizeof(float) * n);
            h2d(m1.p(), sizeof(float) * n);
            h2d(m2.p(), sizeof(float) * n);
            h2d(steps.p(), sizeof(uint32_t) * n);
            h2d(state.p(), sizeof(TrainState));
        }
        fill_f32(stream, grads.p(), n, 0.0f);
        bind(stream);
        NLOD_HIP(hipStreamSynchronize(stream));
    }

    // tcnn's weight file (Trainer::serialize without the optimizer,
    // trainer.h:267-278, written as JSON text by write_weights_to_disk,
    // neural_lod_learning_ui.cu:916-924): nlohmann's text form of a binary
    // value is {"bytes": [...], "subtype": null} (gpu_memory_json.h:36-71);
    // keys sorted as nlohmann's object map sorts them.  The payload is the
    // inference parameters in the network precision, as the reference's
    // Trainer<float, T, T> stores its T copy: "float" for FP32 models, "__half"
    // (each value rounded to nearest even) for FP16 models
    // (neural_lod_learning.h:297 uses T = __half).
    std::string weights_json() {
        NLOD_HIP(hipStreamSynchronize(stream));
        const bool half = model.precision == PREC_FP16;
        std::vector<uint8_t> bytes(sizeof(float) * n);
        NLOD_HIP(hipMemcpy(bytes.data(), opt.ema ? ema.p() : params.p(), bytes.size(), hipMemcpyDeviceToHost));
        if (half) {
            for (size_t k = 0; k < n; ++k) {
                float f;
                std::memcpy(&f, bytes.data() + 4 * k, 4);
                const _Float16 h = (_Float16)f;
                std::memcpy(bytes.data() + 2 * k, &h, 2);
            }
            bytes.resize(2 * n);
        }
        std::string out;
        out.reserve(bytes.size() * 4 + 128);
        out += "{\"n_params\":" + std::to_string(n) + ",\"params_binary\":{\"bytes\":[";
        char num[4];
        for (size_t i = 0; i < bytes.size(); ++i) {
            if (i) out += ',';
            const unsigned v = bytes[i];
            int k = 0;
            if (v >= 100) num[k++] = (char)('0' + v / 100);
            if (v >= 10) num[k++] = (char)('0' + (v / 10) % 10);
            num[k++] = (char)('0' + v % 10);
            out.append(num, (size_t)k);
        }
        out += half ? "],\"subtype\":null},\"params_type\":\"__half\"}" : "],\"subtype\":null},\"params_type\":\"float\"}";
        return out;
    }

    // Trainer::deserialize (trainer.h:280-307): params_type "float" or "__half";
    // params_binary as the {"bytes": [...]} object of the text form.  Other
    // keys (an "optimizer" state) are ignored.
    void load_weights_json(hipStream_t caller, const char* js, size_t len) {
        const std::string text(js, len);
        auto value_of = [&](const char* key) -> size_t {  // index just past the key's ':'
            const std::string k = std::string("\"") + key + "\"";
            const size_t at = text.find(k);
            if (at == std::string::npos) return std::string::npos;
            const size_t colon = text.find(':', at + k.size());
            return colon == std::string::npos ? colon : colon + 1;
        };
        std::string type = "float";
        const size_t tv = value_of("params_type");
        if (tv != std::string::npos) {
            const size_t q0 = text.find('"', tv), q1 = q0 == std::string::npos ? q0 : text.find('"', q0 + 1);
            if (q1 == std::string::npos) fail(NLOD_ERR_INVALID_ARGUMENT, "weights JSON: bad params_type");
            type = text.substr(q0 + 1, q1 - q0 - 1);
        }
        if (type != "float" && type != "__half")
            fail(NLOD_ERR_INVALID_ARGUMENT, "weights JSON: params_type must be float or __half");
        const size_t bv = value_of("params_binary");
        if (bv == std::string::npos) fail(NLOD_ERR_INVALID_ARGUMENT, "weights JSON: no params_binary");
        const size_t bytes_key = text.find("\"bytes\"", bv);
        const size_t open = bytes_key == std::string::npos ? bytes_key : text.find('[', bytes_key);
        if (open == std::string::npos) fail(NLOD_ERR_INVALID_ARGUMENT, "weights JSON: params_binary has no bytes");
        std::vector<uint8_t> bytes;
        bytes.reserve(sizeof(float) * n);
        size_t i = open + 1;
        for (;;) {
            while (i < len && (text[i] == ' ' || text[i] == '\n' || text[i] == '\t' || text[i] == '\r' || text[i] == ','))
                ++i;
            if (i >= len) fail(NLOD_ERR_INVALID_ARGUMENT, "weights JSON: unterminated bytes");
            if (text[i] == ']') break;
            unsigned v = 0;
            int digits = 0;
            while (i < len && text[i] >= '0' && text[i] <= '9' && digits < 4) v = v * 10 + (unsigned)(text[i++] - '0'), ++digits;
            if (digits == 0 || v > 255) fail(NLOD_ERR_INVALID_ARGUMENT, "weights JSON: bytes must be 0..255");
            bytes.push_back((uint8_t)v);
        }
        const size_t elem = type == "float" ? 4 : 2;
        if (bytes.size() != elem * n)
            fail(NLOD_ERR_INVALID_ARGUMENT, "weights JSON holds " + std::to_string(bytes.size() / elem) +
                                                " parameters, the model has " + std::to_string(n));
        std::vector<float> host(n);
        if (elem == 4) {
            std::memcpy(host.data(), bytes.data(), bytes.size());
        } else {
            for (size_t k = 0; k < n; ++k) {
                _Float16 h;
                std::memcpy(&h, bytes.data() + 2 * k, 2);
                host[k] = (float)h;
            }
        }
        DevBuf tmp;
        tmp.ensure(sizeof(float) * n);
        NLOD_HIP(hipMemcpy(tmp.p, host.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        set_params(caller, tmp.as<float>());
        NLOD_HIP(hipStreamSynchronize(caller));  // tmp is released on return
    }
};

}  // namespace nlod

// ============================================================== C ABI
struct nlod_model;
struct nlod_trainer {
    std::unique_ptr<nlod::Trainer> t;
};

namespace nlod {
Model& model_of(nlod_model* m);
}

extern "C" {

int nlod_trainer_create(nlod_model* model, const char* optimizer_json, const char* loss_json,
                        const char* loss_scale_json, uint32_t seed, uint32_t flags, nlod_trainer** out) {
    return nlod::guarded([&] {
        NLOD_CHECK(model && out && optimizer_json && loss_json, "null argument");
        *out = nullptr;
        NLOD_CHECK((flags & ~3u) == 0, "unknown trainer flags");
        const nlod::Json opt = nlod::Json::parse(optimizer_json);
        const nlod::Json loss = nlod::Json::parse(loss_json);
        const nlod::Json scaler = nlod::Json::parse(loss_scale_json ? loss_scale_json : "{}");
        auto h = std::make_unique<nlod_trainer>();
        h->t = std::make_unique<nlod::Trainer>(nlod::model_of(model), opt, loss, scaler, seed, (flags & 1u) != 0,
                                               (flags & 2u) == 0);
        *out = h.release();
    });
}

int nlod_trainer_destroy(nlod_trainer* tr) {
    return nlod::guarded([&] { delete tr; });
}

int nlod_trainer_step(nlod_trainer* tr, void* stream, uint32_t n, const float* d_in, const float* d_target) {
    return nlod::guarded([&] {
        NLOD_CHECK(tr, "null trainer");
        tr->t->step(nlod::as_stream(stream), n, d_in, d_target);
    });
}

int nlod_trainer_get_status(nlod_trainer* tr, nlod_trainer_status* out) {
    return nlod::guarded([&] {
        NLOD_CHECK(tr && out, "null argument");
        const nlod::TrainState s = tr->t->read_state();
        std::memset(out, 0, sizeof(*out));
        out->loss = s.loss;
        out->grad_norm = std::sqrt(s.grad_sq);
        out->loss_scale = s.loss_scale;
        out->learning_rate = s.lr;
        out->step = s.step;
        out->skipped_steps = s.skipped;
        out->last_step_skipped = s.found_inf;
    });
}

int nlod_trainer_params(nlod_trainer* tr, float** d_params, float** d_inference_params, float** d_gradients) {
    return nlod::guarded([&] {
        NLOD_CHECK(tr, "null trainer");
        nlod::Trainer& t = *tr->t;
        if (d_params) *d_params = t.params.p();
        if (d_inference_params) *d_inference_params = t.opt.ema ? t.ema.p() : t.params.p();
        if (d_gradients) *d_gradients = t.grads.p();
    });
}

int nlod_trainer_copy(nlod_trainer* tr, void* stream, int which, float* d_dst) {
    return nlod::guarded([&] {
        NLOD_CHECK(tr && d_dst, "null argument");
        nlod::Trainer& t = *tr->t;
        const float* src = nullptr;
        switch (which) {
            case 0: src = t.params.p(); break;
            case 1: src = t.opt.ema ? t.ema.p() : t.params.p(); break;
            case 2: src = t.grads.p(); break;
            case 3: src = t.m1.p(); break;
            case 4: src = t.m2.p(); break;
            default: nlod::fail(NLOD_ERR_INVALID_ARGUMENT, "which must be 0..4");
        }
        hipStream_t s = nlod::as_stream(stream);
        NLOD_HIP(hipEventRecord(t.ev_out, t.stream));
        NLOD_HIP(hipStreamWaitEvent(s, t.ev_out, 0));
        NLOD_HIP(hipMemcpyAsync(d_dst, src, sizeof(float) * t.n, hipMemcpyDeviceToDevice, s));
    });
}

int nlod_trainer_set_params(nlod_trainer* tr, void* stream, const float* d_params) {
    return nlod::guarded([&] {
        NLOD_CHECK(tr && d_params, "null argument");
        tr->t->set_params(nlod::as_stream(stream), d_params);
    });
}

int nlod_trainer_serialize(nlod_trainer* tr, int with_optimizer_state, void* buf, size_t buf_len, size_t* needed) {
    return nlod::guarded([&] {
        NLOD_CHECK(tr, "null trainer");
        const size_t b = tr->t->snapshot_bytes(with_optimizer_state != 0);
        if (needed) *needed = b;
        if (!buf) return;
        NLOD_CHECK(buf_len >= b, "snapshot buffer too small");
        tr->t->serialize(with_optimizer_state != 0, static_cast<uint8_t*>(buf));
    });
}

int nlod_trainer_deserialize(nlod_trainer* tr, const void* buf, size_t len) {
    return nlod::guarded([&] {
        NLOD_CHECK(tr && buf, "null argument");
        tr->t->drop_graphs();
        tr->t->deserialize(static_cast<const uint8_t*>(buf), len);
    });
}

int nlod_trainer_weights_json(nlod_trainer* tr, char* buf, size_t buf_len, size_t* needed) {
    return nlod::guarded([&] {
        NLOD_CHECK(tr, "null trainer");
        const std::string js = tr->t->weights_json();
        if (needed) *needed = js.size() + 1;
        if (!buf) return;
        NLOD_CHECK(buf_len >= js.size() + 1, "weights JSON buffer too small");
        std::memcpy(buf, js.c_str(), js.size() + 1);
    });
}

int nlod_trainer_load_weights_json(nlod_trainer* tr, void* stream, const char* json, size_t len) {
    return nlod::guarded([&] {
        NLOD_CHECK(tr && json, "null argument");
        tr->t->drop_graphs();
        tr->t->load_weights_json(nlod::as_stream(stream), json, len);
    });
}

int nlod_loss_evaluate(void* stream, int loss_type, uint64_t n_elements, float loss_scale, const float* d_pred,
                       const float* d_target, float* d_values, float* d_gradients) {
    return nlod::guarded([&] {
        NLOD_CHECK(loss_type >= nlod::LOSS_L2 && loss_type <= nlod::LOSS_BCE, "invalid loss type");
        NLOD_CHECK(n_elements == 0 || (d_pred && d_target && d_gradients), "null argument");
        nlod::loss_evaluate(nlod::as_stream(stream), loss_type, n_elements, loss_scale, d_pred, d_target, d_values,
                            d_gradients);
    });
}

int nlod_adam_step(void* stream, const nlod_adam_config* cfg, uint64_t n, uint64_t n_matrix, float loss_scale,
                   uint32_t optimizer_step, float* d_weights, float* d_gradients, float* d_first_moments,
                   float* d_second_moments, uint32_t* d_param_steps) {
    return nlod::guarded([&] {
        NLOD_CHECK(cfg, "null config");
        NLOD_CHECK(n == 0 || (d_weights && d_gradients && d_first_moments && d_second_moments && d_param_steps),
                   "null argument");
        NLOD_CHECK(n_matrix <= n && loss_scale > 0.0f && optimizer_step > 0, "invalid sizes, scale or step");
        hipStream_t s = nlod::as_stream(stream);
        nlod::AdamParams a{};
        a.beta1 = cfg->beta1;
        a.beta2 = cfg->beta2;
        a.epsilon = cfg->epsilon;
        a.l2_reg = cfg->l2_reg;
        a.relative_decay = cfg->relative_decay;
        a.absolute_decay = cfg->absolute_decay;
        a.clipping = cfg->clipping_magnitude;
        a.non_matrix_lr_factor = cfg->non_matrix_learning_rate_factor;
        a.optimize_matrix = cfg->optimize_matrix_params;
        a.optimize_non_matrix = cfg->optimize_non_matrix_params;
        a.adabound = cfg->adabound;
        // a one-off state: this step's scale, learning rate and optimizer step count
        nlod::DevBuf st;
        st.ensure(sizeof(nlod::TrainState));
        nlod::TrainState h{};
        h.loss_scale = h.applied_scale = loss_scale;
        h.lr = cfg->learning_rate;
        h.lr_factor = 1.0f;
        h.step = optimizer_step;
        NLOD_HIP(hipMemcpyAsync(st.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
        nlod::adam_update(s, a, n, n_matrix, st.as<nlod::TrainState>(), d_weights, nullptr, d_gradients,
                          d_first_moments, d_second_moments, d_param_steps, false);
        NLOD_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
