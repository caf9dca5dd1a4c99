"""GPU parity of the encodings and the network (libnlod through the C ABI) vs
the FP64 CPU oracle on identical parameters and inputs.

Tolerances (stated per the path's precision policy):
  FP32 form: |gpu - oracle| <= 1e-5 * max(|oracle|, max over the batch of
             |oracle| in the same output column) -- 1e-5 relative to the
             column's scale (degree-8 SH polynomials lose ~3e-6 absolute to
             cancellation in any FP32 evaluation)
  FP16 MFMA form: |gpu - oracle| <= 2e-3 (absolute, per output)
"""
import numpy as np
import pytest
import torch

from neural_lod_amd import configs, synthetic
from oracle import model as M

pytestmark = pytest.mark.gpu


def _inputs(cfg, n, seed):
    if cfg["n_input_dims"] == 6:
        return synthetic.visibility_queries(n, seed)
    if cfg["n_input_dims"] == 10:
        return synthetic.throughput_queries(n, seed=seed)
    if cfg["n_input_dims"] == 9:
        return synthetic.throughput_queries(n, seed=seed)[:, :9].copy()
    return np.random.default_rng(seed).uniform(0, 1, (n, cfg["n_input_dims"])).astype(np.float32)


def _model(cfg, precision=0):
    from neural_lod_amd.tcnn import NetworkWithInputEncoding
    return NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"],
                                    precision)


def _params(cfg, m, seed, grid_range):
    L = M.Layout(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"])
    p = M.xavier_uniform(L, seed, grid_range=grid_range)
    assert p.size == m.n_params
    return L, p


def _fp32_close(got, ref, what):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64).reshape(got.shape)
    col = np.abs(ref).max(axis=0, keepdims=True) if ref.ndim == 2 else np.abs(ref).max()
    tol = 1e-5 * np.maximum(np.abs(ref), col) + 1e-30
    bad = np.abs(got - ref) > tol
    assert not bad.any(), f"{what}: {bad.sum()} of {bad.size} outside FP32 tolerance, max err {np.abs(got - ref).max():.3e}"


GRID_VARIANTS = {
    "dense_smoothstep": {"n_input_dims": 3, "n_output_dims": 2,
                         "encoding": {"otype": "DenseGrid", "n_levels": 5, "n_features_per_level": 4,
                                      "base_resolution": 4, "per_level_scale": 1.6, "interpolation": "Smoothstep"},
                         "network": {"otype": "FullyFusedMLP", "n_neurons": 32, "n_hidden_layers": 2,
                                     "output_activation": "None"}},
    "tiled_nearest_f1": {"n_input_dims": 3, "n_output_dims": 1,
                         "encoding": {"otype": "TiledGrid", "n_levels": 8, "n_features_per_level": 1,
                                      "base_resolution": 6, "per_level_scale": 1.4, "interpolation": "Nearest"},
                         "network": {"otype": "FullyFusedMLP", "n_neurons": 16, "n_hidden_layers": 1}},
    "hash_f8": {"n_input_dims": 3, "n_output_dims": 1,
                "encoding": {"otype": "HashGrid", "n_levels": 4, "n_features_per_level": 8, "log2_hashmap_size": 12,
                             "base_resolution": 8, "per_level_scale": 2.0},
                "network": {"otype": "FullyFusedMLP", "n_neurons": 64, "n_hidden_layers": 3}},
}
ALL_CFGS = dict(configs.ALL, **GRID_VARIANTS)


@pytest.mark.parametrize("name", sorted(ALL_CFGS))
def test_encoding_fp32(cuda, name):
    cfg = ALL_CFGS[name]
    print(f"case encoding {name}", flush=True)
    m = _model(cfg)
    L, p = _params(cfg, m, 1, grid_range=1.0)
    x = _inputs(cfg, 3000, 2)
    ref = M.encode(L, p, x)  # [N][K]
    pt = torch.from_numpy(p).cuda()
    m.set_params(pt)
    enc = m.encode(torch.from_numpy(x).cuda()).cpu().numpy()  # [K][N]
    torch.cuda.synchronize()
    _fp32_close(enc.T, ref, f"encoding {name}")


@pytest.mark.parametrize("name", sorted(ALL_CFGS))
def test_inference_fp32(cuda, name):
    cfg = ALL_CFGS[name]
    print(f"case inference fp32 {name}", flush=True)
    m = _model(cfg)
    L, p = _params(cfg, m, 3, grid_range=1.0)
    for n in (1, 127, 1000, 4099):
        x = _inputs(cfg, n, 4)
        ref = M.forward(L, p, x)[:, :cfg["n_output_dims"]]
        m.set_params(torch.from_numpy(p).cuda())
        out = m.inference(torch.from_numpy(x).cuda()).cpu().numpy()
        _fp32_close(out, ref, f"inference {name} n={n}")


FP16_CASES = ["visibility", "throughput", "reference_visibility", "reference_throughput", "dense_smoothstep",
              "hash_f8"]


@pytest.mark.parametrize("name", FP16_CASES)
@pytest.mark.parametrize("grid_range", [1e-4, 1.0])
def test_inference_fp16_mfma(cuda, name, grid_range):
    """FP16 MFMA form vs the FP64 oracle: max abs error <= 2e-3 of the output
    scale (max(1, max |y|): outputs without a bounded activation grow past 1).
    reference_throughput is the reference's production throughput network
    (two SH degree 8 blocks, K = 160, width 128, linear outputs).
    dense_smoothstep (F = 4, all levels small enough for the LDS pass) and
    hash_f8 (F = 8, 64 KB levels: one in the LDS pass, the rest in the
    per-level pass) cover the other feature widths."""
    cfg = ALL_CFGS[name]
    print(f"case inference fp16 {name} grid_range={grid_range}", flush=True)
    m = _model(cfg, precision=1)
    L, p = _params(cfg, m, 5, grid_range=grid_range)
    for n in (1, 200, 4096, 20000):
        x = _inputs(cfg, n, 6)
        ref = M.forward(L, p, x)[:, :cfg["n_output_dims"]]
        m.set_params(torch.from_numpy(p).cuda())
        out = m.inference(torch.from_numpy(x).cuda()).cpu().numpy()
        err = np.abs(out - ref).max()
        tol = 2e-3 * max(1.0, float(np.abs(ref).max()))
        print(f"  n={n} max abs err {err:.3e} (tolerance {tol:.3e})", flush=True)
        assert err <= tol, f"FP16 MFMA {name} n={n}: max abs error {err:.3e} > {tol:.3e}"


@pytest.mark.parametrize("name", FP16_CASES)
def test_encoding_pass_features_fp16(cuda, name):
    """The FP16 encoding pass feature by feature (coarse levels from the LDS
    pass, the other levels from the per-level rows, SH / identity / pad rows
    from the non-grid row) against the FP64 oracle's encoding of the same
    FP16-rounded tables: within FP16 rounding of each value, and the K padding
    beyond the encoded width exactly 0."""
    cfg = ALL_CFGS[name]
    print(f"case encoding pass features {name}", flush=True)
    m = _model(cfg, precision=1)
    L, p = _params(cfg, m, 11, grid_range=1.0)
    m.set_params(torch.from_numpy(p).cuda())
    n = 3000  # not a multiple of 16: the last tile is partly padding
    x = _inputs(cfg, n, 12)
    kp = (L.net.n_input + 31) // 32 * 32
    buf = m.encode_fragments(torch.from_numpy(x).cuda()).view(torch.float16).cpu().numpy().astype(np.float64)
    tiles = (n + 15) // 16
    got = buf[:tiles * kp * 16].reshape(tiles, kp // 2, 16, 2).transpose(0, 2, 1, 3).reshape(tiles * 16, kp)
    p16 = p.astype(np.float16).astype(np.float32)  # the table copy the encoding pass gathers from
    ref = M.encode(L, p16, x)
    w = ref.shape[1]
    err = np.abs(got[:n, :w] - ref)
    # FP16 rounding of the result, plus the FP32 arithmetic before it: interpolation
    # (1e-6), or the degree-8 SH polynomials, whose terms reach ~30 before they
    # cancel (an FP32 evaluation of them on the CPU is off by up to 4e-6)
    sh8 = any(e.get("otype") == "SphericalHarmonics" and e.get("degree", 4) > 4
              for e in cfg["encoding"].get("nested", [cfg["encoding"]]))
    bound = 2.0 ** -10 * np.abs(ref) + (8e-6 if sh8 else 1e-6)
    assert np.all(err <= bound), f"{name}: {np.sum(err > bound)} features outside FP16 rounding, max {err.max():.3e}"
    assert not np.any(got[:n, w:]), f"{name}: K padding rows are not 0"
    assert not np.any(got[n:]), f"{name}: padding queries of the last tile are not 0"


def test_fp16_matches_fp32_form(cuda):
    cfg = configs.VISIBILITY
    m16, m32 = _model(cfg, 1), _model(cfg, 0)
    _, p = _params(cfg, m32, 9, grid_range=1.0)
    pt = torch.from_numpy(p).cuda()
    m16.set_params(pt)
    m32.set_params(pt)
    x = torch.from_numpy(_inputs(cfg, 1 << 16, 10)).cuda()
    a = m16.inference(x)
    b = m32.inference(x)
    assert (a - b).abs().max().item() <= 2e-3


@pytest.mark.parametrize("name", ["visibility", "throughput", "reference_visibility", "dense_smoothstep"])
def test_backward_fp32(cuda, name):
    cfg = ALL_CFGS[name]
    print(f"case backward fp32 {name}", flush=True)
    m = _model(cfg)
    L, p = _params(cfg, m, 11, grid_range=0.5)
    n = 2000
    x = _inputs(cfg, n, 12)
    rng = np.random.default_rng(13)
    d_out = np.zeros((n, L.net.n_output_padded))
    d_out[:, :cfg["n_output_dims"]] = rng.normal(size=(n, cfg["n_output_dims"]))
    g_ref, _ = M.backward(L, p, x, d_out)
    m.set_params(torch.from_numpy(p).cuda())
    xt = torch.from_numpy(x).cuda()
    out, ctx = m.forward(xt)
    ref_out = M.forward(L, p, x)[:, :cfg["n_output_dims"]]
    _fp32_close(out.cpu().numpy(), ref_out, "forward")
    d_params, _ = m.backward(ctx, xt, torch.from_numpy(d_out[:, :cfg["n_output_dims"]].astype(np.float32)).cuda())
    g = d_params.cpu().numpy().astype(np.float64)
    scale = np.abs(g_ref).max()
    err = np.abs(g - g_ref)
    # FP32 accumulation over a batch of 2000 and float atomics: relative to the gradient scale
    assert err.max() <= 2e-5 * scale, f"{name}: max grad err {err.max():.3e} vs scale {scale:.3e}"
    assert np.linalg.norm(g - g_ref) <= 1e-5 * np.linalg.norm(g_ref)


def test_backward_accumulate_and_ignore(cuda):
    cfg = configs.VISIBILITY
    m = _model(cfg)
    _, p = _params(cfg, m, 14, grid_range=0.5)
    m.set_params(torch.from_numpy(p).cuda())
    x = torch.from_numpy(_inputs(cfg, 512, 15)).cuda()
    d = torch.randn(512, 1, device="cuda")
    _, ctx = m.forward(x)
    g1, _ = m.backward(ctx, x, d)
    g2 = g1.clone()
    m.backward(ctx, x, d, d_params=g2, mode=2)
    assert torch.allclose(g2, 2 * g1, rtol=1e-5, atol=1e-9)
    g3 = torch.full_like(g1, 7.0)
    m.backward(ctx, x, d, d_params=g3, mode=0)
    assert (g3 == 7.0).all()


def test_encode_backward_grid(cuda):
    cfg = configs.VISIBILITY
    m = _model(cfg)
    L, p = _params(cfg, m, 16, grid_range=1.0)
    n = 5000
    x = _inputs(cfg, n, 17)
    d_enc = np.random.default_rng(18).normal(size=(n, L.enc_width))
    g_ref = np.zeros(L.n_params)
    import ctypes as C
    M._lib().nlo_encode_backward(L.segs, len(L.segs), L.enc_width, M._fp(x), x.shape[1], n, M._dp(d_enc),
                                 M._dp(g_ref))
    m.set_params(torch.from_numpy(p).cuda())
    g = m.encode_backward(torch.from_numpy(x).cuda(),
                          torch.from_numpy(np.ascontiguousarray(d_enc.T.astype(np.float32))).cuda()).cpu().numpy()
    assert np.abs(g[:L.n_net_params]).max() == 0
    err = np.abs(g - g_ref).max()
    assert err <= 1e-5 * np.abs(g_ref).max() + 1e-6, err


def test_sh_input_gradient(cuda):
    """dL/dinput through SH + identity (the encodings that support it)."""
    cfg = {"n_input_dims": 4, "n_output_dims": 1,
           "encoding": {"otype": "Composite", "nested": [
               {"otype": "SphericalHarmonics", "n_dims_to_encode": 3, "degree": 5},
               {"otype": "Identity", "n_dims_to_encode": 1, "scale": 2.0, "offset": -1.0}]},
           "network": {"otype": "FullyFusedMLP", "n_neurons": 32, "n_hidden_layers": 2, "output_activation": "None"}}
    m = _model(cfg)
    L, p = _params(cfg, m, 19, grid_range=0.0)
    rng = np.random.default_rng(20)
    x = rng.uniform(0.1, 0.9, (64, 4)).astype(np.float32)
    d_out = rng.normal(size=(64, 1)).astype(np.float32)
    m.set_params(torch.from_numpy(p).cuda())
    xt = torch.from_numpy(x).cuda()
    _, ctx = m.forward(xt, prepare_input_gradients=True)
    _, d_in = m.backward(ctx, xt, torch.from_numpy(d_out).cuda(), input_gradients=True)
    d_in = d_in.cpu().numpy()
    # chain rule with the oracle's analytic dL/d(enc) and central differences of
    # the (smooth) encoding alone, so no ReLU kink is straddled
    d_o = np.zeros((64, 16))
    d_o[:, :1] = d_out
    _, d_enc = M.backward(L, p, x, d_o)
    h = 1e-3
    for k in range(4):
        xp, xm = x.copy(), x.copy()
        xp[:, k] += h
        xm[:, k] -= h
        de = (M.encode(L, p, xp) - M.encode(L, p, xm)) / (xp[:, k:k + 1].astype(np.float64) - xm[:, k:k + 1])
        expect = (de * d_enc).sum(axis=1)
        assert np.allclose(d_in[:, k], expect, rtol=1e-3, atol=1e-5), k


def test_torch_module_gradcheck(cuda):
    """torch.autograd.Function harness: gradients vs central finite differences
    (h = 1e-3) of the GPU forward itself, and vs the FP64 oracle gradient."""
    from neural_lod_amd.tcnn import Module
    cfg = configs.VISIBILITY
    mod = Module(6, 1, cfg["encoding"], cfg["network"], seed=21, loss_scale=128.0).cuda()
    L = M.Layout(6, 1, cfg["encoding"], cfg["network"])
    with torch.no_grad():
        mod.params.copy_(torch.from_numpy(M.xavier_uniform(L, 21, grid_range=0.5)))
    x = torch.from_numpy(_inputs(cfg, 256, 22)).cuda()
    w = torch.from_numpy(np.random.default_rng(24).normal(size=(256, 1)).astype(np.float32)).cuda()
    loss = (mod(x) * w).sum()
    loss.backward()
    g = mod.params.grad.detach().cpu().numpy().astype(np.float64)
    d_out = np.zeros((256, 16))
    d_out[:, :1] = w.cpu().numpy()
    g_ref, _ = M.backward(L, mod.params.detach().cpu().numpy(), x.cpu().numpy(), d_out)
    assert np.abs(g - g_ref).max() <= 2e-5 * np.abs(g_ref).max()
    rng = np.random.default_rng(23)
    grid_g = np.abs(g_ref[L.n_net_params:])
    big_grid = L.n_net_params + np.argsort(-grid_g)[:500]  # touched table entries with large gradient
    idx = np.concatenate([rng.choice(L.n_net_params, 20, replace=False), rng.choice(big_grid, 20, replace=False)])

    def fd(i, h):
        p0 = mod.params[i].item()
        mod.params[i] = p0 + h
        lp = (mod(x) * w).sum().item()
        mod.params[i] = p0 - h
        lm = (mod(x) * w).sum().item()
        mod.params[i] = p0
        return (lp - lm) / (2 * h)

    checked = 0
    with torch.no_grad():
        for i in idx:
            tol = 2e-3 * max(1.0, abs(g[i])) + 5e-3 * abs(g[i])
            f1, f2 = fd(i, 1e-3), fd(i, 5e-4)
            if abs(f1 - f2) > tol:
                continue  # a ReLU kink inside [p - h, p + h]: central differences are not valid there
            checked += 1
            assert abs(f1 - g[i]) <= tol, (i, f1, f2, g[i])
    assert checked >= 30, f"only {checked} of {len(idx)} parameters away from kinks"


def test_invalid_calls(cuda):
    from neural_lod_amd._lib import NlodError
    cfg = configs.VISIBILITY
    m = _model(cfg)
    x = torch.rand(16, 6, device="cuda")
    with pytest.raises(NlodError, match="parameters are not set"):
        m.inference(x)
    with pytest.raises(ValueError):
        m.set_params(torch.zeros(10, device="cuda"))
