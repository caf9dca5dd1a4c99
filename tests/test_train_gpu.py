"""GPU parity of the training step: loss kernels and Adam against the oracle,
the FP32 trainer against the oracle's training loop, EMA / decay schedule,
dynamic loss scaling, hipGraph replay, snapshots, and the FP16 loss curve
against the FP32 plain form (within 2% at equal steps).

Tolerances: loss values / gradients 1e-5 relative (FP32 vs FP64); Adam 1e-6
relative on the weights (float32 both sides; powf may differ by an ulp), the
moments within one rounding of their O(1) terms (FMA contraction); trainer loss
curve vs oracle 1e-5 at step 0, 2e-3 after (Adam's sign-like first steps
amplify last-bit gradient differences on near-zero gradients).
"""
import numpy as np
import pytest
import torch

from neural_lod_amd import configs, synthetic
from oracle import model as M

pytestmark = pytest.mark.gpu

SMALL_VIS = {
    "n_input_dims": 6, "n_output_dims": 1,
    "encoding": {"otype": "Composite", "nested": [
        {"otype": "HashGrid", "n_dims_to_encode": 3, "n_levels": 4, "n_features_per_level": 2,
         "log2_hashmap_size": 12, "base_resolution": 8, "per_level_scale": 1.5},
        {"otype": "SphericalHarmonics", "n_dims_to_encode": 3, "degree": 4}]},
    "network": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "Sigmoid",
                "n_neurons": 32, "n_hidden_layers": 2},
}
ADAM = {"otype": "Adam", "learning_rate": 1e-2, "beta1": 0.9, "beta2": 0.99, "epsilon": 1e-8, "l2_reg": 1e-8}
BCE = {"otype": "BinaryCrossEntropy"}


def _model(cfg, precision):
    from neural_lod_amd.tcnn import NetworkWithInputEncoding
    return NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"],
                                    precision)


def _trainer(cfg, precision, optimizer=ADAM, loss=BCE, seed=1337, scale=None, graphs=False, fused=True):
    from neural_lod_amd.tcnn import Trainer
    m = _model(cfg, precision)
    return Trainer(m, optimizer, loss, seed=seed, loss_scale_config=scale, graphs=graphs, fused=fused)


def _batches(n, count, seed):
    out = []
    for k in range(count):
        x = synthetic.visibility_queries(n, seed + k)
        out.append((x, synthetic.sphere_visibility_labels(x)))
    return out


@pytest.mark.parametrize("kind,code", [("L2", 0), ("RelativeL2", 1), ("BinaryCrossEntropy", 2)])
def test_loss_kernels_vs_oracle(cuda, kind, code):
    from neural_lod_amd.tcnn import loss_evaluate
    rng = np.random.default_rng(5)
    n, dims = 1000, 3
    pred = rng.uniform(0.0, 1.0, (n, dims)).astype(np.float32)
    pred[:4, 0] = [0.0, 1.0, 1e-7, 1.0 - 1e-7]  # BCE clamp edges
    target = (rng.uniform(size=(n, dims)) < 0.4).astype(np.float32) if code == 2 else \
        rng.uniform(0, 2, (n, dims)).astype(np.float32)
    vals_ref, g_ref = M.loss(kind, pred.astype(np.float64), target, dims, loss_scale=128.0)
    v, g = loss_evaluate(code, torch.from_numpy(pred).cuda(), torch.from_numpy(target).cuda(), 128.0)
    np.testing.assert_allclose(v.cpu().numpy(), vals_ref, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(g.cpu().numpy(), g_ref, rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("extra", [{}, {"non_matrix_learning_rate_factor": 0.5, "relative_decay": 0.1},
                                   {"adabound": True, "clipping_magnitude": 0.05},
                                   {"optimize_non_matrix_params": False}])
def test_adam_kernel_vs_oracle(cuda, extra):
    from neural_lod_amd.tcnn import adam_step
    rng = np.random.default_rng(6)
    n, n_matrix = 10007, 1001  # ragged tail and a matrix/grid boundary inside a 4-wide group
    w0 = rng.uniform(-0.3, 0.3, n).astype(np.float32)
    cfg = dict(ADAM, **extra)
    w, m, v = w0.copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)
    steps = np.zeros(n, np.uint32)
    wt, mt, vt = (torch.from_numpy(a.copy()).cuda() for a in (w0, m, v))
    st = torch.zeros(n, dtype=torch.int32, device="cuda")
    for s in range(1, 4):
        g = rng.normal(size=n).astype(np.float32) * 64.0
        g[rng.uniform(size=n) < 0.3] = 0.0
        M.adam_step(w, g, m, v, steps, n_matrix, 64.0, cfg["learning_rate"], cfg["beta1"], cfg["beta2"],
                    cfg["epsilon"], cfg["l2_reg"], opt_step=s, **extra)
        adam_step(cfg, wt, torch.from_numpy(g).cuda(), mt, vt, st, n_matrix, 64.0, s)
    assert np.array_equal(st.cpu().numpy().astype(np.uint32), steps)
    np.testing.assert_allclose(wt.cpu().numpy(), w, rtol=1e-6, atol=1e-8)
    # moments: the GPU contracts b*m + (1-b)*g into one FMA (as nvcc does for the
    # reference), the oracle rounds twice: one rounding of the O(1) terms apart
    np.testing.assert_allclose(mt.cpu().numpy(), m, rtol=1e-6, atol=2e-7)
    np.testing.assert_allclose(vt.cpu().numpy(), v, rtol=1e-6, atol=2e-7)


def test_trainer_fp32_vs_oracle_loop(cuda):
    cfg = SMALL_VIS
    tr = _trainer(cfg, 0, scale={"initial_scale": 128.0, "dynamic": False})
    L = M.Layout(6, 1, cfg["encoding"], cfg["network"])
    p0 = tr.params().cpu().numpy()
    assert np.array_equal(p0, tr.model.initial_params(1337))
    batches = _batches(256, 8, 40)
    losses_ref, w_ref, _ = M.train(L, p0, batches, ADAM, "BinaryCrossEntropy", loss_scale=128.0)
    losses = []
    for x, t in batches:
        tr.training_step(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
        losses.append(tr.loss())
    losses = np.array(losses)
    print("gpu", losses, "\noracle", losses_ref)
    assert abs(losses[0] - losses_ref[0]) <= 1e-5 * losses_ref[0]
    np.testing.assert_allclose(losses, losses_ref, rtol=2e-3)
    w = tr.params().cpu().numpy()
    close = np.abs(w - w_ref) <= 1e-4 + 1e-3 * np.abs(w_ref)
    assert close.mean() >= 0.99, f"{(~close).sum()} of {w.size} parameters differ"
    st = tr.status()
    assert st["step"] == 8 and st["skipped_steps"] == 0 and st["loss_scale"] == 128.0


def test_trainer_ema_and_exponential_decay(cuda):
    cfg = SMALL_VIS
    opt = {"otype": "Ema", "decay": 0.9, "nested": {"otype": "ExponentialDecay", "decay_start": 2,
                                                    "decay_interval": 2, "decay_base": 0.5, "nested": ADAM}}
    tr = _trainer(cfg, 0, optimizer=opt, scale={"initial_scale": 1.0, "dynamic": False})
    L = M.Layout(6, 1, cfg["encoding"], cfg["network"])
    p0 = tr.params().cpu().numpy()
    batches = _batches(256, 6, 50)
    losses_ref, w_ref, ema_ref = M.train(L, p0, batches, opt, "BinaryCrossEntropy", loss_scale=1.0)
    for x, t in batches:
        tr.training_step(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    st = tr.status()
    factors = M.exp_decay_factors(6, 0.5, 2, 2, 10000000)
    assert st["learning_rate"] == pytest.approx(1e-2 * factors[-1], rel=1e-6)
    ema = tr.inference_params().cpu().numpy()
    close = np.abs(ema - ema_ref) <= 1e-4 + 1e-3 * np.abs(ema_ref)
    assert close.mean() >= 0.99
    # the model's inference path reads the EMA weights
    x = torch.from_numpy(batches[0][0]).cuda()
    out = tr.model.inference(x).cpu().numpy()
    ref = M.forward(L, ema, batches[0][0])[:, :1]
    assert np.abs(out - ref).max() <= 1e-5


def test_dynamic_loss_scale_skips_overflow(cuda):
    """FP16: a huge scale overflows the FP16 deltas; the step is skipped, the
    scale backs off and the parameters stay untouched until a finite step."""
    cfg = SMALL_VIS
    tr = _trainer(cfg, 1, scale={"initial_scale": float(2 ** 30)})
    p0 = tr.params()
    x, t = _batches(1024, 1, 60)[0]
    xt, tt = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    tr.training_step(xt, tt)
    st = tr.status()
    assert st["last_step_skipped"] == 1 and st["skipped_steps"] == 1 and st["step"] == 0
    assert st["loss_scale"] == float(2 ** 29)
    assert torch.equal(tr.params(), p0)
    for _ in range(40):
        tr.training_step(xt, tt)
        if tr.status()["step"] > 0:
            break
    st = tr.status()
    assert st["step"] == 1 and st["last_step_skipped"] == 0
    assert st["loss_scale"] == float(2 ** 30) / 2 ** st["skipped_steps"]
    assert not torch.equal(tr.params(), p0)
    assert torch.isfinite(tr.gradients()).all(), "a finite step leaves no trace of the skipped ones"
    assert torch.isfinite(tr.params()).all()


def test_loss_scale_growth(cuda):
    tr = _trainer(SMALL_VIS, 0, scale={"initial_scale": 1024.0, "growth_interval": 3})
    x, t = _batches(512, 1, 61)[0]
    xt, tt = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    scales = []
    for _ in range(7):
        tr.training_step(xt, tt)
        scales.append(tr.status()["loss_scale"])
    assert scales == [1024.0, 1024.0, 2048.0, 2048.0, 2048.0, 4096.0, 4096.0]


@pytest.mark.parametrize("precision", [0, 1])
def test_trainer_graph_replay_matches_eager(cuda, precision):
    cfg = SMALL_VIS
    eager = _trainer(cfg, precision, seed=7)
    graph = _trainer(cfg, precision, seed=7, graphs=True)
    sizes = [512, 512, 300, 512, 512, 300, 300, 512]
    for k, n in enumerate(sizes):
        x, t = _batches(n, 1, 70 + k)[0]
        xt, tt = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
        eager.training_step(xt, tt)
        graph.training_step(xt, tt)
        a, b = eager.status(), graph.status()  # eager host work between replays
        assert b["loss"] == pytest.approx(a["loss"], rel=1e-5, abs=1e-9), k
        assert b["step"] == a["step"] == k + 1
    pa, pb = eager.params(), graph.params()
    close = (pa - pb).abs() <= 1e-5 + 1e-4 * pa.abs()
    assert close.float().mean().item() >= 0.999


@pytest.mark.parametrize("precision", [0, 1])
def test_trainer_graph_after_model_buffers_grow(cuda, precision):
    """A captured step holds the addresses of the model's scratch buffers.
    Inference on the same model at larger batches between steps (as the
    reference does between training steps, neural_lod_learning.cu:379,781)
    frees and moves those buffers; the steps after it must not replay the old
    graph.  The graph trainer must track an eager trainer given the same calls."""
    cfg = SMALL_VIS
    eager = _trainer(cfg, precision, seed=11)
    graph = _trainer(cfg, precision, seed=11, graphs=True)
    big = torch.from_numpy(synthetic.visibility_queries(1 << 16, 99)).cuda()
    for k in range(10):
        x, t = _batches(512, 1, 200 + k)[0]
        xt, tt = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
        eager.training_step(xt, tt)
        graph.training_step(xt, tt)
        a, b = eager.status(), graph.status()
        assert b["loss"] == pytest.approx(a["loss"], rel=1e-5, abs=1e-9), k
        if k in (3, 6):  # the graph was captured at step 1 and replayed since
            n = 1 << (13 + k // 3)
            ye = eager.model.inference(big[:n])
            yg = graph.model.inference(big[:n])
            assert torch.equal(ye, yg) or (ye - yg).abs().max().item() <= 1e-5
    pa, pb = eager.params(), graph.params()
    close = (pa - pb).abs() <= 1e-5 + 1e-4 * pa.abs()
    assert close.float().mean().item() >= 0.999


@pytest.mark.parametrize("name", ["visibility", "throughput", "small_vis"])
def test_fused_training_pass_matches_unfused(cuda, name):
    """FP16: the fused forward + loss + backward + dW kernel gives the loss and
    gradients of the separate kernels (same FP16 rounding points; dW summed in
    another order)."""
    cfg = SMALL_VIS if name == "small_vis" else configs.ALL[name]
    n = 3000
    if cfg["n_input_dims"] == 6:
        x = synthetic.visibility_queries(n, 90)
        t = synthetic.sphere_visibility_labels(x)
        loss = BCE
    else:
        x = synthetic.throughput_queries(n, seed=91)
        t = np.stack([0.3 + 0.2 * np.sin(4 * x[:, 0]), 0.5 + 0.1 * x[:, 1], 0.2 + 0.3 * x[:, 2]], 1).astype(np.float32)
        loss = {"otype": "RelativeL2"}
    scale = {"initial_scale": 1024.0, "dynamic": False}
    tf = _trainer(cfg, 1, loss=loss, seed=5, scale=scale)
    tu = _trainer(cfg, 1, loss=loss, seed=5, scale=scale, fused=False)
    xt, tt = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    tf.training_step(xt, tt)
    tu.training_step(xt, tt)
    lf, lu = tf.loss(), tu.loss()
    assert lf == pytest.approx(lu, rel=1e-6), (lf, lu)
    gf, gu = tf.gradients().cpu().numpy().astype(np.float64), tu.gradients().cpu().numpy().astype(np.float64)
    L = M.Layout(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"])
    from test_backward_fp16_gpu import _blocks
    for blk, sl in _blocks(L):
        rel = np.linalg.norm(gf[sl] - gu[sl]) / max(np.linalg.norm(gu[sl]), 1e-30)
        print(f"  {blk}: fused vs unfused rel L2 {rel:.2e}")
        assert rel <= 1e-4, (blk, rel)
    assert torch.equal(tf.params(), tu.params()) or \
        ((tf.params() - tu.params()).abs() > 1e-6).float().mean().item() < 1e-3


def test_snapshot_roundtrip(cuda):
    cfg = SMALL_VIS
    a = _trainer(cfg, 0, seed=3)
    batches = _batches(256, 5, 80)
    for x, t in batches[:3]:
        a.training_step(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    blob = a.serialize(with_optimizer_state=True)
    b = _trainer(cfg, 0, seed=99)
    b.deserialize(blob)
    assert torch.equal(a.params(), b.params())
    for (m1, m2) in zip(a.moments(), b.moments()):
        assert torch.equal(m1, m2)
    assert a.status()["step"] == b.status()["step"] == 3
    for x, t in batches[3:]:
        for tr in (a, b):
            tr.training_step(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    pa, pb = a.params(), b.params()
    assert ((pa - pb).abs() <= 1e-6 + 1e-5 * pa.abs()).float().mean().item() >= 0.999
    small = a.serialize(False)
    assert len(small) == 24 + 4 * a.model.n_params


def test_fp16_loss_curve_within_2pct_of_fp32(cuda):
    """Visibility network (16-level hash grid + SH4, MLP 64x4) trained on
    analytic sphere-occlusion labels: FP16-MFMA and FP32 plain forms from the
    same initial parameters, for three initialisations; the mean over them of
    the loss of the last 50 of 300 steps agrees within 2%.  One initialisation
    alone is not enough at this size: two runs of one form from different
    seeds end up to ~2% apart, so single pairs measure that spread, not the
    arithmetic."""
    cfg = configs.VISIBILITY
    batches = []
    for k in range(300):
        x = synthetic.visibility_queries(1 << 14, 1000 + k)
        batches.append((torch.from_numpy(x).cuda(), torch.from_numpy(synthetic.sphere_visibility_labels(x)).cuda()))
    finals = {0: [], 1: []}
    for seed in (1337, 7, 2024):
        for prec in (0, 1):
            tr = _trainer(cfg, prec, seed=seed)
            losses = []
            for x, t in batches:
                tr.training_step(x, t)
                losses.append(tr.loss())
            losses = np.array(losses)
            assert losses[-50:].mean() < 0.6 * losses[:10].mean(), "the task must be learnable"
            finals[prec].append(losses[-50:].mean())
            print(f"seed {seed} precision {prec}: {losses[::60]} -> {finals[prec][-1]:.5f}", flush=True)
    l32, l16 = np.mean(finals[0]), np.mean(finals[1])
    assert abs(l16 - l32) <= 0.02 * l32, f"FP16 {l16:.5f} vs FP32 {l32:.5f}"


def test_fp16_loss_curve_mesh_labels(cuda):
    """Same comparison on mesh visibility labels (rays from the occupied 64^3
    voxels of a blob mesh, label = nearest hit inside the voxel), batches
    drawn on the GPU by the labeler: mean loss of the last 50 of 300 steps of
    the FP16-MFMA form within 2% of the FP32 plain form."""
    from neural_lod_amd.labels import VisibilityLabeler
    cfg = configs.VISIBILITY
    lb = VisibilityLabeler(torch.from_numpy(synthetic.blob_mesh(50_000, seed=0)).cuda(), 64)
    batches = [lb.sample(1 << 14, seed=7, first=k << 14) for k in range(300)]
    curves = {}
    for prec in (0, 1):
        tr = _trainer(cfg, prec, seed=1337)
        losses = []
        for x, t in batches:
            tr.training_step(x, t)
            losses.append(tr.loss())
        curves[prec] = np.array(losses)
        print(prec, curves[prec][::30])
    l32, l16 = curves[0][-50:].mean(), curves[1][-50:].mean()
    assert l32 < 0.8 * curves[0][:10].mean(), "the task must be learnable"
    assert abs(l16 - l32) <= 0.02 * l32, f"FP16 {l16:.5f} vs FP32 {l32:.5f}"


def test_throughput_loss_curve_mc_targets(cuda):
    """Throughput network (hash grid + 2 x SH4 + LoD scalar, MLP 64x5, exp
    outputs) on Monte-Carlo throughput targets of a blob mesh over a 3-level
    LoD pyramid (valid samples only), RelativeL2: mean loss of the last 50 of
    300 steps of the FP16-MFMA form within 2% of the FP32 plain form."""
    from neural_lod_amd.labels import Labeler
    cfg = configs.THROUGHPUT
    lb = Labeler(torch.from_numpy(synthetic.blob_mesh(50_000, seed=0)).cuda(), 32, 3)
    x, t, v = lb.throughput(1 << 16, seed=3, spp=64, max_depth=3)
    keep = v > 0
    x, t = x[keep].contiguous(), t[keep].contiguous()
    nb = 1 << 12
    assert x.shape[0] >= 8 * nb
    batches = [(x[(k % 8) * nb:(k % 8 + 1) * nb], t[(k % 8) * nb:(k % 8 + 1) * nb]) for k in range(300)]
    curves = {}
    for prec in (0, 1):
        # lr 1e-3: at 1e-2 the exp outputs fall to 0, where RelativeL2 has no gradient
        adam = {"otype": "Adam", "learning_rate": 1e-3, "beta1": 0.9, "beta2": 0.99, "epsilon": 1e-15}
        tr = _trainer(cfg, prec, optimizer=adam, loss={"otype": "RelativeL2"}, seed=1337)
        losses = []
        for xb, tb in batches:
            tr.training_step(xb, tb)
            losses.append(tr.loss())
        curves[prec] = np.array(losses)
        print(prec, curves[prec][::30])
    l32, l16 = curves[0][-50:].mean(), curves[1][-50:].mean()
    assert l32 < 0.2 * curves[0][0], "the task must be learnable"
    assert abs(l16 - l32) <= 0.02 * l32, f"FP16 {l16:.5f} vs FP32 {l32:.5f}"


def test_weights_json_roundtrip(cuda):
    """The reference's weight file: nlohmann's JSON text of Trainer::serialize
    (sorted keys, binary as {"bytes": [...], "subtype": null}); float and
    __half payloads load back; a wrong size is rejected."""
    import json
    from neural_lod_amd._lib import NlodError
    cfg = SMALL_VIS
    tr = _trainer(cfg, 0, seed=3)
    for x, t in _batches(512, 3, 40):
        tr.training_step(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    text = tr.weights_json()
    assert text.startswith('{"n_params":') and text.endswith('"params_type":"float"}')
    d = json.loads(text)
    assert d["params_type"] == "float" and d["params_binary"]["subtype"] is None
    w = np.array(d["params_binary"]["bytes"], np.uint8).view(np.float32)
    assert d["n_params"] == w.size
    assert np.array_equal(w, tr.inference_params().cpu().numpy())
    tr2 = _trainer(cfg, 0, seed=99)
    tr2.load_weights_json(text)
    assert torch.equal(tr2.params(), tr.inference_params())
    h = w.astype(np.float16)
    d16 = {"n_params": int(w.size), "optimizer": {"otype": "Adam"},
           "params_binary": {"bytes": list(h.tobytes()), "subtype": None}, "params_type": "__half"}
    tr2.load_weights_json(json.dumps(d16, indent=1))
    assert np.array_equal(tr2.params().cpu().numpy(), h.astype(np.float32))
    d16["params_binary"]["bytes"] = d16["params_binary"]["bytes"][:-2]
    with pytest.raises(NlodError, match="parameters"):
        tr2.load_weights_json(json.dumps(d16))


def test_weights_json_fp16_model_writes_half(cuda):
    """FP16 models write the reference's Trainer<float, __half, __half> file:
    params_type "__half", 2 bytes per parameter, the inference parameters
    rounded to nearest even; it loads back into an FP16 and an FP32 trainer."""
    import json
    cfg = SMALL_VIS
    tr = _trainer(cfg, 1, seed=4)
    for x, t in _batches(512, 3, 50):
        tr.training_step(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    text = tr.weights_json()
    assert text.endswith('"params_type":"__half"}')
    d = json.loads(text)
    h = np.array(d["params_binary"]["bytes"], np.uint8).view(np.float16)
    assert d["n_params"] == h.size == tr.model.n_params
    ref = tr.inference_params().cpu().numpy().astype(np.float16)
    assert np.array_equal(h.view(np.uint16), ref.view(np.uint16))
    for prec in (1, 0):
        tr2 = _trainer(cfg, prec, seed=98)
        tr2.load_weights_json(text)
        assert np.array_equal(tr2.params().cpu().numpy(), h.astype(np.float32))
