// Wavefront queries of the neural LoD path tracer on gfx950: for each ray
// segment of a wavefront, the visibility network and then the throughput
// network (rptr/cuda/neural-lod/wavefront_neural_throughput_visibility_lod.cu:
// 464-720 runs, per wavefront, input generation, visibility inference, a
// visibility check, throughput input generation, throughput inference and the
// bounce update as separate launches with a host round trip in between).
//
// FP16 models: one persistent launch per wavefront.  Every wave takes 16-query
// tiles; per tile it encodes the visibility inputs straight into MFMA B
// fragments (hash-grid gathers from the FP16 table copy, SH), runs the
// visibility MLP, then the throughput encoding and MLP, all with weights of
// both networks resident in LDS, and writes the visibility and, for segments
// whose visibility reaches the threshold, the throughput (0 otherwise).
// Tiles with no visible segment skip the throughput network.  The threshold is
// the engine's fixed value, or per segment the threshold of the voxel the
// segment's march stopped in (neural_result.voxel_threshold, compared as
// visibility >= threshold in check_visibility, :320-325).
//
// The launch shape (blocks per CU; or the split per-level encode + MLP
// launches, faster for large wavefronts) is chosen per wavefront-size bucket by
// an autotuner that times the candidates once.  A sequence of wavefronts can be
// captured in a hipGraph and replayed.  FP32 models run the plain form:
// separate encode / MLP launches per network and a combine kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "device_common.h"
#include "fused_net.h"
#include "mfma_common.h"
#include "model.h"

namespace nlod {
namespace {

constexpr int kBlock = 256;

template <int W, int F, int MAXD>
__global__ __launch_bounds__(kBlock) void wavefront_kernel(uint32_t N, NetArgs vis, NetArgs thr, float threshold,
                                                           const float* __restrict__ seg_threshold,
                                                           const half_t* __restrict__ packed_vis,
                                                           const half_t* __restrict__ packed_thr,
                                                           const float* __restrict__ vis_in,
                                                           const float* __restrict__ thr_in,
                                                           float* __restrict__ vis_out, float* __restrict__ thr_out) {
    extern __shared__ float4 smem4[];
    half_t* sw = reinterpret_cast<half_t*>(smem4);
    for (uint32_t k = threadIdx.x; k < vis.packed_halves / 8; k += kBlock)
        reinterpret_cast<float4*>(sw + vis.lds_offset)[k] = reinterpret_cast<const float4*>(packed_vis)[k];
    for (uint32_t k = threadIdx.x; k < thr.packed_halves / 8; k += kBlock)
        reinterpret_cast<float4*>(sw + thr.lds_offset)[k] = reinterpret_cast<const float4*>(packed_thr)[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const uint32_t n_tiles = (N + 15) / 16;
    const uint32_t n_waves = gridDim.x * (kBlock / 64);
    for (uint32_t tile = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); tile < n_tiles; tile += n_waves) {
        const uint32_t q = tile * 16 + n;
        const bool valid = q < N;
        const uint32_t qc = valid ? q : N - 1;  // a clamped duplicate keeps loads in bounds; not stored
        const floatx4 ov = run_net<W, F, MAXD>(vis, sw, vis_in + (size_t)qc * vis.n_in);
        // visibility of query n = output 0, held by lane group 0
        const float v0 = act_fwd(vis.out_act, ov[0]);
        const float vis_n = __shfl(v0, n, 64);
        if (valid) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t oi = 4 * g + i;
                if (oi < vis.n_out) vis_out[(size_t)q * vis.n_out + oi] = act_fwd(vis.out_act, ov[i]);
            }
        }
        const float t_q = seg_threshold ? seg_threshold[qc] : threshold;
        const bool hit = valid && vis_n >= t_q;
        if (__ballot(hit) == 0) {  // no visible segment in the tile: no throughput query
            if (valid && g == 0)
                for (uint32_t oi = 0; oi < thr.n_out; ++oi) thr_out[(size_t)q * thr.n_out + oi] = 0.0f;
            continue;
        }
        const floatx4 ot = run_net<W, F, MAXD>(thr, sw, thr_in + (size_t)qc * thr.n_in);
        if (valid) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t oi = 4 * g + i;
                if (oi < thr.n_out) thr_out[(size_t)q * thr.n_out + oi] = hit ? act_fwd(thr.out_act, ot[i]) : 0.0f;
            }
        }
    }
}

// Plain form: thr_out rows of segments below their threshold are zeroed.
__global__ void gate_kernel(uint32_t N, uint32_t n_vis, uint32_t n_thr, float threshold,
                            const float* __restrict__ seg_threshold, const float* __restrict__ vis_out,
                            float* __restrict__ thr_out) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N) return;
    const float t_q = seg_threshold ? seg_threshold[q] : threshold;
    if (!(vis_out[(size_t)q * n_vis] >= t_q))
        for (uint32_t i = 0; i < n_thr; ++i) thr_out[(size_t)q * n_thr + i] = 0.0f;
}

// wavefront-size buckets of the autotuner: [2^k, 2^(k+2)) for k = 8, 10, ..., 20
constexpr int kBuckets = 8;
int bucket_of(uint32_t n) {
    int b = 0;
    while (b < kBuckets - 1 && n >= (256u << (2 * (b + 1)))) ++b;
    return b;
}

}  // namespace

struct Wavefront {
    Model& vis;
    Model& thr;
    float threshold;
    bool fp16 = false, fused_ok = false;
    uint32_t F = 2, maxdeg = 4;
    size_t smem = 0;
    // tuned choice per bucket: 0 = split launches, k > 0 = fused kernel with k blocks per CU
    int choice[kBuckets];
    bool tuned = false;
    int max_blocks_per_cu = 1;
    DevBuf tune_in_v, tune_in_t, tune_out_v, tune_out_t;
    // Captured sequences, keyed by counts and buffers, with the tick of their
    // last use (least recently used goes first beyond kMaxGraphs).  Every graph
    // was captured while both models' epochs had the values in graph_stamp: a
    // replay after one of them moved a buffer or rebound its weights would read
    // freed or stale memory, so a changed epoch drops them all.
    static constexpr size_t kMaxGraphs = 16;
    struct CachedGraph {
        hipGraphExec_t exec;
        uint64_t last_use;
    };
    std::map<std::vector<uint32_t>, CachedGraph> graphs;
    uint64_t use_tick = 0;
    std::pair<uint64_t, uint64_t> graph_stamp{0, 0};
    std::pair<uint64_t, uint64_t> stamp() const { return {vis.epoch, thr.epoch}; }
    void drop_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second.exec);
        graphs.clear();
    }

    Wavefront(Model& v, Model& t, float thr_) : vis(v), thr(t), threshold(thr_) {
        if (v.precision != t.precision) fail(NLOD_ERR_INVALID_ARGUMENT, "both networks must use the same precision");
        if (v.L.net.n_output < 1) fail(NLOD_ERR_INVALID_ARGUMENT, "visibility network has no output");
        fp16 = v.precision == PREC_FP16;
        for (int b = 0; b < kBuckets; ++b) choice[b] = 0;
        if (!fp16) return;
        uint32_t Fv, Ft, dv, dt;
        layout_features(v.L, Fv, dv);
        layout_features(t.L, Ft, dt);
        const uint32_t W = v.L.net.width;
        fused_ok = W == t.L.net.width && (W == 32 || W == 64) && Fv == Ft;
        F = Fv;
        maxdeg = std::max(dv, dt);
        smem = mfma_packed_bytes(v.L.net) + mfma_packed_bytes(t.L.net);
        if (smem > 160 * 1024) fused_ok = false;
        if (fused_ok) {
            dispatch_wf(W, F, maxdeg, [&](auto Wc, auto Fc, auto Dc) {
                auto kern = wavefront_kernel<decltype(Wc)::value, decltype(Fc)::value, decltype(Dc)::value>;
                max_blocks_per_cu = std::max(1, occupancy_blocks((const void*)kern, kBlock, smem));
            });
            for (int b = 0; b < kBuckets; ++b) choice[b] = 1;
        }
    }

    ~Wavefront() { drop_graphs(); }

    NetArgs args_of(const Model& m, uint32_t lds_offset) const { return net_args_of(m, lds_offset); }

    void run_fused(hipStream_t s, uint32_t N, const float* vin, const float* tin, float* vout, float* tout,
                   int blocks_per_cu, const float* seg_thr = nullptr) {
        const NetArgs av = args_of(vis, 0);
        const NetArgs at = args_of(thr, av.packed_halves);
        const uint32_t tiles = div_round_up(N, 16);
        const uint32_t blocks =
            std::max<uint32_t>(1, std::min<uint32_t>(div_round_up(tiles, kBlock / 64), (uint32_t)(cu_count() * blocks_per_cu)));
        dispatch_wf(vis.L.net.width, F, maxdeg, [&](auto Wc, auto Fc, auto Dc) {
            auto kern = wavefront_kernel<decltype(Wc)::value, decltype(Fc)::value, decltype(Dc)::value>;
            hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), smem, s, N, av, at, threshold, seg_thr,
                               static_cast<const half_t*>(vis.packed_fwd(true)),
                               static_cast<const half_t*>(thr.packed_fwd(true)), vin, tin, vout, tout);
        });
        NLOD_LAUNCH_CHECK();
    }

    void run_split(hipStream_t s, uint32_t N, const float* vin, const float* tin, float* vout, float* tout,
                   const float* seg_thr = nullptr) {
        vis.inference(s, N, vin, vout);
        thr.inference(s, N, tin, tout);
        hipLaunchKernelGGL(gate_kernel, dim3(div_round_up(N, 256)), dim3(256), 0, s, N, vis.L.net.n_output,
                           thr.L.net.n_output, threshold, seg_thr, vout, tout);
        NLOD_LAUNCH_CHECK();
    }

    // shape < 0: the tuned choice for N; 0: split launches; k > 0: fused kernel, k blocks per CU
    void run(hipStream_t s, uint32_t N, const float* vin, const float* tin, float* vout, float* tout, int shape = -1,
             const float* seg_thr = nullptr) {
        if (N == 0) return;
        vis.require_params(true);
        thr.require_params(true);
        const int c = shape < 0 ? choice[bucket_of(N)] : shape;
        if (c > 0 && !(fp16 && fused_ok)) fail(NLOD_ERR_UNSUPPORTED, "the fused wavefront kernel does not apply here");
        if (c > max_blocks_per_cu) fail(NLOD_ERR_INVALID_ARGUMENT, "more blocks per CU than fit");
        if (c > 0) run_fused(s, N, vin, tin, vout, tout, c, seg_thr);
        else run_split(s, N, vin, tin, vout, tout, seg_thr);
    }

    // Times each candidate once per bucket (synchronizes the stream).
    void tune(hipStream_t s, uint32_t max_n) {
        if (!fp16 || !fused_ok) {
            tuned = true;
            return;
        }
        max_n = std::max<uint32_t>(max_n, 256);
        const uint32_t nv = vis.L.n_input_dims, nt = thr.L.n_input_dims;
        tune_in_v.ensure(sizeof(float) * (size_t)max_n * nv);
        tune_in_t.ensure(sizeof(float) * (size_t)max_n * nt);
        tune_out_v.ensure(sizeof(float) * (size_t)max_n * vis.L.net.n_output);
        tune_out_t.ensure(sizeof(float) * (size_t)max_n * thr.L.net.n_output);
        // inputs in [0, 1) from a fixed pattern (any in-range values exercise the same work)
        std::vector<float> h(std::max<size_t>((size_t)max_n * nv, (size_t)max_n * nt));
        for (size_t i = 0; i < h.size(); ++i) h[i] = (float)((i * 2654435761ull) % 1000003ull) / 1000003.0f;
        NLOD_HIP(hipMemcpy(tune_in_v.p, h.data(), sizeof(float) * (size_t)max_n * nv, hipMemcpyHostToDevice));
        NLOD_HIP(hipMemcpy(tune_in_t.p, h.data(), sizeof(float) * (size_t)max_n * nt, hipMemcpyHostToDevice));
        hipEvent_t e0, e1;
        NLOD_HIP(hipEventCreate(&e0));
        NLOD_HIP(hipEventCreate(&e1));
        for (int b = 0; b < kBuckets; ++b) {
            const uint32_t n = std::min<uint32_t>(max_n, 256u << (2 * b + 1));  // middle of the bucket
            float best = 1e30f;
            int best_c = 1;
            for (int c = 0; c <= max_blocks_per_cu; ++c) {
                auto launch = [&] {
                    if (c == 0) run_split(s, n, tune_in_v.as<float>(), tune_in_t.as<float>(), tune_out_v.as<float>(),
                                          tune_out_t.as<float>());
                    else run_fused(s, n, tune_in_v.as<float>(), tune_in_t.as<float>(), tune_out_v.as<float>(),
                                   tune_out_t.as<float>(), c);
                };
                launch();  // warm
                const int reps = 5;
                NLOD_HIP(hipEventRecord(e0, s));
                for (int r = 0; r < reps; ++r) launch();
                NLOD_HIP(hipEventRecord(e1, s));
                NLOD_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                NLOD_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    best_c = c;
                }
            }
            choice[b] = best_c;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        tuned = true;
        drop_graphs();
    }

    // A sequence of wavefronts with host-known sizes, concatenated in the
    // buffers; optionally captured once per size signature and replayed.
    void run_batches(hipStream_t s, const std::vector<uint32_t>& counts, const float* vin, const float* tin,
                     float* vout, float* tout, bool use_graph, hipStream_t capture_stream,
                     const float* seg_thr = nullptr) {
        auto body = [&](hipStream_t st) {
            size_t off = 0;
            for (uint32_t n : counts) {
                run(st, n, vin + off * vis.L.n_input_dims, tin + off * thr.L.n_input_dims,
                    vout + off * vis.L.net.n_output, tout + off * thr.L.net.n_output, -1,
                    seg_thr ? seg_thr + off : nullptr);
                off += n;
            }
        };
        if (!use_graph) {
            body(s);
            return;
        }
        std::vector<uint32_t> key(counts);
        for (const void* p : {(const void*)vin, (const void*)tin, (const void*)vout, (const void*)tout,
                              (const void*)seg_thr}) {
            const uint64_t v = reinterpret_cast<uintptr_t>(p);
            key.push_back((uint32_t)v);
            key.push_back((uint32_t)(v >> 32));
        }
        if (!graphs.empty() && stamp() != graph_stamp) drop_graphs();
        auto it = graphs.find(key);
        if (it == graphs.end()) {
            body(s);  // eager first: workspaces of the split path get allocated outside capture
            const auto before = stamp();
            hipGraph_t g = nullptr;
            NLOD_HIP(hipStreamBeginCapture(capture_stream, hipStreamCaptureModeRelaxed));
            try {
                body(capture_stream);
            } catch (...) {
                (void)hipStreamEndCapture(capture_stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            NLOD_HIP(hipStreamEndCapture(capture_stream, &g));
            hipGraphExec_t exec = nullptr;
            const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            NLOD_HIP(e);
            if (stamp() != before) {  // a buffer moved during capture: not kept
                (void)hipGraphExecDestroy(exec);
                return;
            }
            if (graphs.empty()) graph_stamp = before;
            if (graphs.size() >= kMaxGraphs) {
                auto lru = graphs.begin();
                for (auto j = graphs.begin(); j != graphs.end(); ++j)
                    if (j->second.last_use < lru->second.last_use) lru = j;
                (void)hipGraphExecDestroy(lru->second.exec);
                graphs.erase(lru);
            }
            graphs.emplace(key, CachedGraph{exec, ++use_tick});
            return;  // this call's results came from the eager run
        }
        it->second.last_use = ++use_tick;
        NLOD_HIP(hipGraphLaunch(it->second.exec, s));
    }
};

}  // namespace nlod

// ============================================================== C ABI
struct nlod_model;
struct nlod_wavefront {
    std::unique_ptr<nlod::Wavefront> w;
    hipStream_t capture = nullptr;
    hipEvent_t ev = nullptr;
};

namespace nlod {
Model& model_of(nlod_model* m);
}

extern "C" {

int nlod_wavefront_create(nlod_model* visibility, nlod_model* throughput, float threshold, nlod_wavefront** out) {
    return nlod::guarded([&] {
        NLOD_CHECK(visibility && throughput && out, "null argument");
        *out = nullptr;
        auto h = std::make_unique<nlod_wavefront>();
        h->w = std::make_unique<nlod::Wavefront>(nlod::model_of(visibility), nlod::model_of(throughput), threshold);
        *out = h.release();
    });
}

int nlod_wavefront_destroy(nlod_wavefront* wf) {
    return nlod::guarded([&] {
        if (!wf) return;
        if (wf->capture) (void)hipStreamDestroy(wf->capture);
        if (wf->ev) (void)hipEventDestroy(wf->ev);
        delete wf;
    });
}

int nlod_wavefront_tune(nlod_wavefront* wf, void* stream, uint32_t max_n) {
    return nlod::guarded([&] {
        NLOD_CHECK(wf, "null argument");
        wf->w->tune(nlod::as_stream(stream), max_n);
    });
}

int nlod_wavefront_choice(const nlod_wavefront* wf, uint32_t n, int* choice_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(wf && choice_out, "null argument");
        *choice_out = (wf->w->fp16 && wf->w->fused_ok) ? wf->w->choice[nlod::bucket_of(n)] : 0;
    });
}

int nlod_wavefront_run(nlod_wavefront* wf, void* stream, uint32_t n, const float* d_vis_in, const float* d_thr_in,
                       float* d_vis_out, float* d_thr_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(wf && (n == 0 || (d_vis_in && d_thr_in && d_vis_out && d_thr_out)), "null argument");
        wf->w->run(nlod::as_stream(stream), n, d_vis_in, d_thr_in, d_vis_out, d_thr_out);
    });
}

int nlod_wavefront_run_shape(nlod_wavefront* wf, void* stream, uint32_t n, const float* d_vis_in,
                             const float* d_thr_in, float* d_vis_out, float* d_thr_out, int shape) {
    return nlod::guarded([&] {
        NLOD_CHECK(wf && (n == 0 || (d_vis_in && d_thr_in && d_vis_out && d_thr_out)), "null argument");
        NLOD_CHECK(shape >= 0, "shape must be 0 (split) or a block count per CU");
        wf->w->run(nlod::as_stream(stream), n, d_vis_in, d_thr_in, d_vis_out, d_thr_out, shape);
    });
}

int nlod_wavefront_run_thresholds(nlod_wavefront* wf, void* stream, uint32_t n, const float* d_vis_in,
                                  const float* d_thr_in, const float* d_thresholds, float* d_vis_out,
                                  float* d_thr_out, int shape) {
    return nlod::guarded([&] {
        NLOD_CHECK(wf && (n == 0 || (d_vis_in && d_thr_in && d_thresholds && d_vis_out && d_thr_out)),
                   "null argument");
        NLOD_CHECK(shape >= -1, "shape must be -1 (engine's choice), 0 (split) or a block count per CU");
        wf->w->run(nlod::as_stream(stream), n, d_vis_in, d_thr_in, d_vis_out, d_thr_out, shape, d_thresholds);
    });
}

int nlod_wavefront_max_shape(const nlod_wavefront* wf, int* max_shape_out) {
    return nlod::guarded([&] {
        NLOD_CHECK(wf && max_shape_out, "null argument");
        *max_shape_out = (wf->w->fp16 && wf->w->fused_ok) ? wf->w->max_blocks_per_cu : 0;
    });
}

namespace {
int run_batches_abi(nlod_wavefront* wf, void* stream, uint32_t n_batches, const uint32_t* counts,
                    const float* d_vis_in, const float* d_thr_in, const float* d_thresholds, float* d_vis_out,
                    float* d_thr_out, int use_graph, bool need_thresholds) {
    return nlod::guarded([&] {
        NLOD_CHECK(wf && (n_batches == 0 || counts), "null argument");
        std::vector<uint32_t> c(counts, counts + n_batches);
        uint64_t total = 0;
        for (uint32_t v : c) total += v;
        NLOD_CHECK(total == 0 || (d_vis_in && d_thr_in && d_vis_out && d_thr_out), "null argument");
        NLOD_CHECK(total == 0 || !need_thresholds || d_thresholds, "null argument");
        hipStream_t s = nlod::as_stream(stream);
        if (!use_graph) {
            wf->w->run_batches(s, c, d_vis_in, d_thr_in, d_vis_out, d_thr_out, false, nullptr, d_thresholds);
            return;
        }
        if (!wf->capture) {
            NLOD_HIP(hipStreamCreateWithFlags(&wf->capture, hipStreamNonBlocking));
            NLOD_HIP(hipEventCreateWithFlags(&wf->ev, hipEventDisableTiming));
        }
        // replay (or capture) on the engine's stream, ordered after and before the caller's work
        NLOD_HIP(hipEventRecord(wf->ev, s));
        NLOD_HIP(hipStreamWaitEvent(wf->capture, wf->ev, 0));
        wf->w->run_batches(wf->capture, c, d_vis_in, d_thr_in, d_vis_out, d_thr_out, true, wf->capture,
                           d_thresholds);
        NLOD_HIP(hipEventRecord(wf->ev, wf->capture));
        NLOD_HIP(hipStreamWaitEvent(s, wf->ev, 0));
    });
}
}  // namespace

int nlod_wavefront_run_batches(nlod_wavefront* wf, void* stream, uint32_t n_batches, const uint32_t* counts,
                               const float* d_vis_in, const float* d_thr_in, float* d_vis_out, float* d_thr_out,
                               int use_graph) {
    return run_batches_abi(wf, stream, n_batches, counts, d_vis_in, d_thr_in, nullptr, d_vis_out, d_thr_out,
                           use_graph, false);
}

int nlod_wavefront_run_batches_thresholds(nlod_wavefront* wf, void* stream, uint32_t n_batches,
                                          const uint32_t* counts, const float* d_vis_in, const float* d_thr_in,
                                          const float* d_thresholds, float* d_vis_out, float* d_thr_out,
                                          int use_graph) {
    return run_batches_abi(wf, stream, n_batches, counts, d_vis_in, d_thr_in, d_thresholds, d_vis_out, d_thr_out,
                           use_graph, true);
}

}  // extern "C"
