"""ORACLE — test infrastructure only (never imported by the product).

Layout restatement of tiny-cuda-nn's encoding tree and network (widths,
padding, parameter order) plus ctypes access to the FP64 operators in
model_oracle.c.  Paths are relative to rptr/ext/tiny-cuda-nn/.

Layout rules restated here:
  * NetworkWithInputEncoding (include/tiny-cuda-nn/network_with_input_encoding.h:44-51):
    encoding created with alignment 8 (src/encoding.cu:85,155-157), then
    set_alignment(16) for FullyFusedMLP (src/network.cu:111-130);
  * Encoding::set_alignment (include/tiny-cuda-nn/encoding.h:83-85);
  * Composite: nested created with alignment 1, then the inter-encoding
    alignment fix (encodings/composite.h:189-199); extra padding goes to the
    last nested encoding (composite.h:381-386);
  * RepeatedComposite (neural_lod): one shared instance per nested encoding,
    repeated n_repetitions times, padding split per repetition
    (encodings/repeated_composite.h:293-323);
  * grid pads with 0 after its features, SH pads with 1 before its features,
    identity pads with 1 after (grid.h:1041-1051, spherical_harmonics.h:58-63,
    identity.h:63-66);
  * parameters: network first (W0 [W][K], hidden [W][W] x (L-1),
    W_out [pad16(out)][W], row-major), then the encoding's parameters in
    nested order (network_with_input_encoding.h:113-128, fully_fused_mlp.cu:635-678).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import oracle as _vox

SEG_GRID, SEG_SH, SEG_IDENTITY = 0, 1, 2
GRID_TYPES = {"hash": 0, "dense": 1, "tiled": 2}
INTERP = {"linear": 0, "nearest": 1, "smoothstep": 2}
ACT = {"none": 0, "relu": 1, "sigmoid": 2, "exponential": 3}
LOSS = {"binarycrossentropy": 0, "relativel2": 1, "l2": 2}


class Seg(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("in_offset", C.c_int32), ("n_in", C.c_int32),
        ("out_offset", C.c_int32), ("n_out", C.c_int32), ("n_pad", C.c_int32),
        ("n_levels", C.c_int32), ("n_features", C.c_int32), ("log2_hashmap_size", C.c_int32),
        ("base_resolution", C.c_int32), ("per_level_scale", C.c_float),
        ("grid_type", C.c_int32), ("interp", C.c_int32), ("param_offset", C.c_int64),
        ("degree", C.c_int32), ("scale", C.c_float), ("offset", C.c_float),
    ]


class Net(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("n_hidden_layers", C.c_int32), ("n_input", C.c_int32),
        ("n_output", C.c_int32), ("n_output_padded", C.c_int32),
        ("activation", C.c_int32), ("output_activation", C.c_int32),
    ]


def _next_multiple(x, m):
    return (x + m - 1) // m * m


class _Leaf:
    def __init__(self, cfg, n_dims):
        self.cfg = cfg
        self.n_dims = n_dims
        self.pad = 0
        ot = cfg.get("otype", "OneBlob").lower()
        if ot in ("hashgrid", "grid", "densegrid", "tiledgrid"):
            self.kind = SEG_GRID
            default_type = "tiled" if ot == "tiledgrid" else ("dense" if ot == "densegrid" else "hash")
            self.grid_type = GRID_TYPES[cfg.get("type", default_type).lower()]
            self.F = int(cfg.get("n_features_per_level", 2))
            self.n_levels = int(cfg.get("n_levels", 16))
            self.log2T = int(cfg.get("log2_hashmap_size", 19))
            self.base = int(cfg.get("base_resolution", 16))
            self.scale = float(np.float32(cfg.get("per_level_scale", 2.0)))
            self.interp = INTERP[cfg.get("interpolation", "Linear").lower()]
            self.n_out = self.F * self.n_levels
            seg = self.segment(0, 0, 0)
            self.n_params = int(_lib().nlo_grid_levels(C.byref(seg), None, None, None))
        elif ot == "sphericalharmonics":
            self.kind = SEG_SH
            self.degree = int(cfg.get("degree", 4))
            self.n_out = self.degree ** 2
            self.n_params = 0
        elif ot == "identity":
            self.kind = SEG_IDENTITY
            self.n_out = n_dims
            self.id_scale = float(cfg.get("scale", 1.0))
            self.id_offset = float(cfg.get("offset", 0.0))
            self.n_params = 0
        else:
            raise ValueError(f"unsupported encoding {ot}")

    def required_output_alignment(self):
        return self.F if self.kind == SEG_GRID else 1

    def padded_output_width(self):
        return self.n_out + self.pad

    def set_padded_output_width(self, w):
        assert w >= self.n_out
        self.pad = w - self.n_out

    def segment(self, in_off, out_off, param_off):
        s = Seg()
        s.type, s.in_offset, s.n_in = self.kind, in_off, self.n_dims
        s.out_offset, s.n_out, s.n_pad = out_off, self.n_out, self.pad
        if self.kind == SEG_GRID:
            s.n_levels, s.n_features, s.log2_hashmap_size = self.n_levels, self.F, self.log2T
            s.base_resolution, s.per_level_scale = self.base, self.scale
            s.grid_type, s.interp, s.param_offset = self.grid_type, self.interp, param_off
        elif self.kind == SEG_SH:
            s.degree = self.degree
        else:
            s.scale, s.offset = self.id_scale, self.id_offset
        return s

    def flatten(self, in_off, out_off, param_off, segs):
        segs.append(self.segment(in_off, out_off, param_off))
        return out_off + self.padded_output_width()


class _Composite:
    def __init__(self, cfg, n_dims, reps=1):
        nested = cfg["nested"]
        self.reps = reps
        total = sum(int(n.get("n_dims_to_encode", 0)) for n in nested) * reps
        unspecified = n_dims - total
        self.children, self.begins = [], []
        off = 0
        for rep in range(reps):
            for n in nested:
                if "n_dims_to_encode" in n:
                    nd = int(n["n_dims_to_encode"])
                else:
                    nd, unspecified = unspecified, None
                if nd > 0:
                    if rep == 0:
                        self.children.append(_make(n, nd, 1))
                    self.begins.append(off)
                off += nd
        if reps > 1 and total != n_dims:
            raise ValueError("RepeatedComposite must encode exactly its input dims")
        so_far = 0
        for i in range(len(self.children) - 1):
            desired = self.children[i + 1].required_output_alignment()
            need = _next_multiple(so_far + self.children[i].padded_output_width(), desired) - so_far
            self.children[i].set_padded_output_width(need)
            so_far += self.children[i].padded_output_width()
        self.n_params = sum(c.n_params for c in self.children)

    def required_output_alignment(self):
        a = 1
        for c in self.children:
            a = a * c.required_output_alignment() // math.gcd(a, c.required_output_alignment())
        return a

    def padded_output_width(self):
        return sum(c.padded_output_width() for c in self.children) * self.reps

    def set_padded_output_width(self, w):
        prev = self.padded_output_width() - self.reps * self.children[-1].padded_output_width()
        assert w >= prev and (w - prev) % self.reps == 0
        self.children[-1].set_padded_output_width((w - prev) // self.reps)

    def flatten(self, in_off, out_off, param_off, segs):
        k = 0
        for rep in range(self.reps):
            p = param_off
            for c in self.children:
                out_off = c.flatten(in_off + self.begins[k], out_off, p, segs)
                p += c.n_params
                k += 1
        return out_off


def _make(cfg, n_dims, alignment):
    ot = cfg.get("otype", "").lower()
    if ot == "composite":
        e = _Composite(cfg, n_dims)
    elif ot == "repeatedcomposite":
        e = _Composite(cfg, n_dims, int(cfg.get("n_repetitions", 1)))
    else:
        e = _Leaf(cfg, n_dims)
    if alignment > 0:
        a = alignment * e.required_output_alignment() // math.gcd(alignment, e.required_output_alignment())
        e.set_padded_output_width(_next_multiple(e.padded_output_width(), a))
    return e


class Layout:
    """Encoding segments + network shape of a NetworkWithInputEncoding."""

    def __init__(self, n_input_dims, n_output_dims, encoding, network):
        enc = _make(encoding, n_input_dims, 8)
        net_align = 16
        a = net_align * enc.required_output_alignment() // math.gcd(net_align, enc.required_output_alignment())
        enc.set_padded_output_width(_next_multiple(enc.padded_output_width(), a))
        self.n_input_dims = n_input_dims
        self.enc_width = enc.padded_output_width()
        self.net = Net()
        self.net.width = int(network.get("n_neurons", 128))
        self.net.n_hidden_layers = int(network.get("n_hidden_layers", 5))
        self.net.n_input = self.enc_width
        self.net.n_output = n_output_dims
        self.net.n_output_padded = _next_multiple(n_output_dims, 16)
        self.net.activation = ACT[network.get("activation", "ReLU").lower()]
        self.net.output_activation = ACT[network.get("output_activation", "None").lower()]
        self.n_net_params = int(_lib().nlo_net_n_params(C.byref(self.net)))
        segs = []
        enc.flatten(0, 0, self.n_net_params, segs)
        self.segs = (Seg * len(segs))(*segs)
        self.n_params = self.n_net_params + enc.n_params

    def grid_levels(self, seg_index):
        s = self.segs[seg_index]
        offs = np.zeros(s.n_levels + 1, np.uint32)
        sc = np.zeros(s.n_levels, np.float32)
        res = np.zeros(s.n_levels, np.uint32)
        _lib().nlo_grid_levels(C.byref(s), _u32p(offs), _fp(sc), _u32p(res))
        return offs, sc, res


_L = None


def _lib():
    global _L
    if _L is None:
        L = _vox.lib()
        sp, np_ = C.POINTER(Seg), C.POINTER(Net)
        fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.nlo_grid_levels.restype = C.c_int64
        L.nlo_grid_levels.argtypes = [sp, C.c_void_p, C.c_void_p, C.c_void_p]
        L.nlo_net_n_params.restype = C.c_int64
        L.nlo_net_n_params.argtypes = [np_]
        L.nlo_encode.argtypes = [sp, C.c_int, C.c_int, fp, fp, C.c_int, C.c_int64, dp]
        L.nlo_mlp_forward.argtypes = [np_, fp, dp, C.c_int64, dp, C.c_void_p]
        L.nlo_mlp_backward.argtypes = [np_, fp, dp, dp, dp, dp, C.c_int64, dp, dp]
        L.nlo_encode_backward.argtypes = [sp, C.c_int, C.c_int, fp, C.c_int, C.c_int64, dp, dp]
        L.nlo_grid_corners.argtypes = [sp, fp, C.POINTER(C.c_int64)]
        L.nlo_infer_f32.argtypes = [sp, C.c_int, np_, fp, fp, C.c_int, C.c_int64, fp]
        L.nlo_loss.argtypes = [C.c_int, dp, fp, C.c_int64, C.c_int, C.c_int, C.c_double, dp, dp]
        L.nlo_adam_step.argtypes = [C.c_int64, C.c_int64, C.c_float, C.POINTER(AdamCfg), C.c_uint32, fp, fp, fp, fp,
                                    C.POINTER(C.c_uint32)]
        L.nlo_ema_step.argtypes = [C.c_int64, C.c_float, C.c_uint32, fp, fp]
        L.nlo_omp_threads.restype = C.c_int
        _L = L
    return _L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u32p(a):
    return a.ctypes.data_as(C.c_void_p)


def encode(layout: Layout, params: np.ndarray, inputs: np.ndarray) -> np.ndarray:
    """Encoded features, double [N][enc_width]."""
    inputs = np.ascontiguousarray(inputs, np.float32)
    params = np.ascontiguousarray(params, np.float32)
    N = inputs.shape[0]
    enc = np.zeros((N, layout.enc_width), np.float64)
    _lib().nlo_encode(layout.segs, len(layout.segs), layout.enc_width, _fp(params), _fp(inputs), inputs.shape[1], N,
                      _dp(enc))
    return enc


def forward(layout: Layout, params: np.ndarray, inputs: np.ndarray, return_hidden: bool = False):
    """FP64 forward: padded outputs double [N][pad16(out)] (+ encoded, hidden)."""
    params = np.ascontiguousarray(params, np.float32)
    enc = encode(layout, params, inputs)
    N = enc.shape[0]
    n = layout.net
    out = np.zeros((N, n.n_output_padded), np.float64)
    hidden = np.zeros((N, n.n_hidden_layers, n.width), np.float64)
    _lib().nlo_mlp_forward(C.byref(n), _fp(params), _dp(enc), N, _dp(out), hidden.ctypes.data_as(C.c_void_p))
    if return_hidden:
        return out, enc, hidden
    return out


def backward(layout: Layout, params: np.ndarray, inputs: np.ndarray, d_out: np.ndarray):
    """FP64 gradients: (dL/dparams double [n_params], dL/d(enc) [N][enc_width])
    for dL/d(padded output) d_out [N][pad16(out)]."""
    params = np.ascontiguousarray(params, np.float32)
    inputs = np.ascontiguousarray(inputs, np.float32)
    out, enc, hidden = forward(layout, params, inputs, return_hidden=True)
    N = enc.shape[0]
    d_out = np.ascontiguousarray(d_out, np.float64)
    g = np.zeros(layout.n_params, np.float64)
    d_enc = np.zeros_like(enc)
    L = _lib()
    L.nlo_mlp_backward(C.byref(layout.net), _fp(params), _dp(enc), _dp(hidden), _dp(out), _dp(d_out), N, _dp(g),
                       _dp(d_enc))
    g_net = g[: layout.n_net_params].copy()
    g[:] = 0.0
    L.nlo_encode_backward(layout.segs, len(layout.segs), layout.enc_width, _fp(inputs), inputs.shape[1], N,
                          _dp(d_enc), _dp(g))
    g[: layout.n_net_params] = g_net
    return g, d_enc


def infer_f32(layout: Layout, params: np.ndarray, inputs: np.ndarray) -> np.ndarray:
    """The plain FP32 C loop (CPU baseline): float [N][n_output]."""
    inputs = np.ascontiguousarray(inputs, np.float32)
    params = np.ascontiguousarray(params, np.float32)
    N = inputs.shape[0]
    out = np.zeros((N, layout.net.n_output), np.float32)
    _lib().nlo_infer_f32(layout.segs, len(layout.segs), C.byref(layout.net), _fp(params), _fp(inputs),
                         inputs.shape[1], N, _fp(out))
    return out


def loss(kind: str, pred: np.ndarray, target: np.ndarray, n_output: int, loss_scale: float = 1.0):
    pred = np.ascontiguousarray(pred, np.float64)
    target = np.ascontiguousarray(target, np.float32)
    N, stride = pred.shape
    values = np.zeros_like(pred)
    grads = np.zeros_like(pred)
    _lib().nlo_loss(LOSS[kind.lower()], _dp(pred), _fp(target), N, stride, n_output, loss_scale, _dp(values),
                    _dp(grads))
    return values, grads


class AdamCfg(C.Structure):
    _fields_ = [(k, C.c_float) for k in ("lr", "beta1", "beta2", "eps", "l2_reg", "relative_decay", "absolute_decay",
                                          "clipping", "non_matrix_lr_factor")] + \
               [(k, C.c_int) for k in ("optimize_matrix", "optimize_non_matrix", "adabound")]


def adam_config(cfg: dict) -> AdamCfg:
    """tcnn Adam JSON keys (adam.h:214-262) with tcnn's defaults."""
    return AdamCfg(lr=cfg.get("learning_rate", 1e-3), beta1=cfg.get("beta1", 0.9), beta2=cfg.get("beta2", 0.999),
                   eps=cfg.get("epsilon", 1e-8), l2_reg=cfg.get("l2_reg", 1e-8),
                   relative_decay=cfg.get("relative_decay", 0.0), absolute_decay=cfg.get("absolute_decay", 0.0),
                   clipping=cfg.get("clipping_magnitude", 0.0),
                   non_matrix_lr_factor=cfg.get("non_matrix_learning_rate_factor", 1.0),
                   optimize_matrix=int(cfg.get("optimize_matrix_params", True)),
                   optimize_non_matrix=int(cfg.get("optimize_non_matrix_params", True)),
                   adabound=int(cfg.get("adabound", False)))


def adam_step(w, grad, m, v, steps, n_matrix, loss_scale, lr, beta1, beta2, eps, l2_reg, opt_step=1, **extra):
    """In-place float32 Adam step (adam.h:48-119); extra = other tcnn Adam keys."""
    cfg = dict(learning_rate=lr, beta1=beta1, beta2=beta2, epsilon=eps, l2_reg=l2_reg, **extra)
    a = adam_config(cfg)
    _lib().nlo_adam_step(w.size, n_matrix, loss_scale, C.byref(a), opt_step, _fp(w), _fp(grad), _fp(m), _fp(v),
                         steps.ctypes.data_as(C.POINTER(C.c_uint32)))


def ema_step(w, ema, decay, t):
    """In-place EMA (ema.h, full precision); t = optimizer steps taken so far."""
    _lib().nlo_ema_step(w.size, decay, t, _fp(w), _fp(ema))


def exp_decay_factors(n_steps: int, base: float, start: int, interval: int, end: int) -> np.ndarray:
    """Learning-rate factor of each step 0..n-1 (ExponentialDecayOptimizer::step,
    exponential_decay.h:60-71): the float factor is multiplied by base before
    the step whose index s satisfies s >= start, (s - start) % interval == 0, s <= end."""
    f = np.float32(1.0)
    out = np.empty(n_steps, np.float32)
    for s in range(n_steps):
        if s >= start and (s - start) % interval == 0 and s <= end:
            f = np.float32(f * np.float32(base))
        out[s] = f
    return out


def unwrap_optimizer(cfg: dict) -> dict:
    """Flattens Ema{ExponentialDecay{Adam}} style JSON into one dict of settings."""
    out = {"adam": {}, "exp_decay": None, "ema": None}
    node = cfg
    while True:
        t = node.get("otype", "Adam").lower()
        if t == "adam":
            out["adam"] = node
            return out
        if t == "exponentialdecay":
            out["exp_decay"] = dict(base=node.get("decay_base", 0.1), start=node.get("decay_start", 10000),
                                    interval=node.get("decay_interval", 10000), end=node.get("decay_end", 10000000))
        elif t == "ema":
            out["ema"] = node.get("decay", 0.99)
        else:
            raise ValueError(f"optimizer {t} not restated")
        node = node.get("nested", {})


def train(layout: Layout, params: np.ndarray, batches, optimizer: dict, loss_kind: str, loss_scale: float = 1.0):
    """The training loop of tcnn's Trainer restated with the FP64 forward /
    backward and the float32 Adam above: for each (inputs, targets) batch,
    forward, loss (scaled by loss_scale), backward, optimizer step.  Returns
    (losses per step, final master params, final inference params)."""
    o = unwrap_optimizer(optimizer)
    a = o["adam"]
    lr0 = a.get("learning_rate", 1e-3)
    w = np.array(params, np.float32)
    n = w.size
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    steps = np.zeros(n, np.uint32)
    ema = np.zeros(n, np.float32) if o["ema"] is not None else None
    decay = o["exp_decay"]
    factors = exp_decay_factors(len(batches), decay["base"], decay["start"], decay["interval"], decay["end"]) \
        if decay else np.ones(len(batches), np.float32)
    losses = []
    for s, (x, t) in enumerate(batches):
        out = forward(layout, w, x)
        vals, g_out = loss(loss_kind, out[:, :layout.net.n_output], t, layout.net.n_output, loss_scale)
        losses.append(vals.sum())
        d = np.zeros((x.shape[0], layout.net.n_output_padded))
        d[:, :layout.net.n_output] = g_out
        g, _ = backward(layout, w, x, d)
        lr = np.float32(np.float32(lr0) * factors[s])
        extra = {k: a[k] for k in ("relative_decay", "absolute_decay", "clipping_magnitude",
                                   "non_matrix_learning_rate_factor", "optimize_matrix_params",
                                   "optimize_non_matrix_params", "adabound") if k in a}
        adam_step(w, g.astype(np.float32), m, v, steps, layout.n_net_params, loss_scale, lr, a.get("beta1", 0.9),
                  a.get("beta2", 0.999), a.get("epsilon", 1e-8), a.get("l2_reg", 1e-8), opt_step=s + 1, **extra)
        if ema is not None:
            ema_step(w, ema, o["ema"], s + 1)
    return np.array(losses), w, (ema if ema is not None else w.copy())


def grid_corners(seg: Seg, x) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    idx = np.zeros(seg.n_levels * 8, np.int64)
    _lib().nlo_grid_corners(C.byref(seg), _fp(x), idx.ctypes.data_as(C.POINTER(C.c_int64)))
    return idx.reshape(seg.n_levels, 8)


def xavier_uniform(layout: Layout, seed: int, grid_range: float = 1e-4) -> np.ndarray:
    """Seeded parameters: Xavier-uniform MLP matrices (incl. padded output rows,
    gpu_matrix.h:291-306), grid tables U[-grid_range, grid_range] (grid.h:1353-1356)."""
    rng = np.random.default_rng(seed)
    n = layout.net
    parts = []
    shapes = [(n.width, n.n_input)] + [(n.width, n.width)] * (n.n_hidden_layers - 1) + [(n.n_output_padded, n.width)]
    for rows, cols in shapes:
        s = math.sqrt(6.0 / (rows + cols))
        parts.append(rng.uniform(-s, s, size=rows * cols))
    parts.append(rng.uniform(-grid_range, grid_range, size=layout.n_params - layout.n_net_params))
    return np.concatenate(parts).astype(np.float32)


# ---------------------------------------------------------------- tcnn initialisation
# Parameter initialisation of tiny-cuda-nn's Trainer, restated with Python
# integers and numpy uint64 arithmetic (wrapping mod 2^64):
#   * Trainer(seed): std::seed_seq{seed}, two words generated, pcg32 seeded
#     with the first (include/tiny-cuda-nn/trainer.h:53-60);
#   * network first (network_with_input_encoding.h:121-127): every weight
#     matrix Xavier-uniform from the host stream, in order, output matrix with
#     its padded rows (src/fully_fused_mlp.cu:863-888, gpu_matrix.h:291-306);
#   * then each grid (composite.h:414-419, repeated_composite.h:348-353; a
#     shared grid once): U[-1e-4, 1e-4] by generate_random_uniform
#     (grid.h:1353-1356, random.h:63-85, common.h:289-293).

_U64 = (1 << 64) - 1
PCG_MULT = 0x5851F42D4C957F2D


class Pcg32:
    """PCG32, XSH-RR output of the 64-bit LCG (O'Neill 2014, pcg-random.org),
    seeded as pcg32(initstate, initseq = 1) in tcnn's pcg32.h."""

    def __init__(self, initstate: int, initseq: int = 1):
        self.state = 0
        self.inc = ((initseq << 1) | 1) & _U64
        self.next_uint()
        self.state = (self.state + initstate) & _U64
        self.next_uint()

    @staticmethod
    def output(old: int) -> int:
        x = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((x >> rot) | (x << ((-rot) & 31))) & 0xFFFFFFFF

    def next_uint(self) -> int:
        old = self.state
        self.state = (old * PCG_MULT + self.inc) & _U64
        return self.output(old)

    def next_float(self) -> np.float32:
        # (u >> 9 | 0x3f800000) read as a float, minus 1: exactly (u >> 9) / 2^23
        return np.float32((self.next_uint() >> 9) * 2.0 ** -23)

    def jump(self, delta: int):
        """(A, C) with state_{k+delta} = A * state_k + C (mod 2^64), by repeated
        squaring of the affine map (Brown 1994)."""
        acc_a, acc_c, a, c = 1, 0, PCG_MULT, self.inc
        while delta:
            if delta & 1:
                acc_a, acc_c = (acc_a * a) & _U64, (acc_c * a + c) & _U64
            a, c = (a * a) & _U64, ((a + 1) * c) & _U64
            delta >>= 1
        return acc_a, acc_c

    def advance(self, delta: int):
        a, c = self.jump(delta)
        self.state = (a * self.state + c) & _U64

    def next_uints(self, n: int) -> np.ndarray:
        """The next n outputs as uint32, states built by doubling jumps."""
        states = np.empty(max(n, 1), np.uint64)
        states[0] = self.state
        m = 1
        while m < n:
            a, c = self.jump(m)
            k = min(m, n - m)
            states[m:m + k] = states[:k] * np.uint64(a) + np.uint64(c)
            m *= 2
        states = states[:n]
        self.advance(n)
        x = (((states >> np.uint64(18)) ^ states) >> np.uint64(27)) & np.uint64(0xFFFFFFFF)
        rot = states >> np.uint64(59)
        out = (x >> rot) | (x << ((np.uint64(32) - rot) & np.uint64(31)))
        return (out & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def seed_seq_generate(seeds, n: int) -> list:
    """std::seed_seq(seeds).generate over n words, as the C++ standard specifies
    it ([rand.util.seedseq])."""
    v = [int(s) & 0xFFFFFFFF for s in seeds]
    s = len(v)
    b = [0x8B8B8B8B] * n
    t = 11 if n >= 623 else 7 if n >= 68 else 5 if n >= 39 else 3 if n >= 7 else (n - 1) // 2
    p = (n - t) // 2
    q = p + t
    m = max(s + 1, n)
    T = lambda x: x ^ (x >> 27)
    M32 = 0xFFFFFFFF
    for k in range(m):
        r1 = (1664525 * T(b[k % n] ^ b[(k + p) % n] ^ b[(k - 1) % n])) & M32
        r2 = (r1 + (s if k == 0 else (k % n + v[k - 1]) if k <= s else k % n)) & M32
        b[(k + p) % n] = (b[(k + p) % n] + r1) & M32
        b[(k + q) % n] = (b[(k + q) % n] + r2) & M32
        b[k % n] = r2
    for k in range(m, m + n):
        r3 = (1566083941 * T((b[k % n] + b[(k + p) % n] + b[(k - 1) % n]) & M32)) & M32
        r4 = (r3 - k % n) & M32
        b[(k + p) % n] ^= r3
        b[(k + q) % n] ^= r4
        b[k % n] = r4
    return b


def uniform_table(rng: Pcg32, n: int, lo: float, hi: float) -> np.ndarray:
    """generate_random_uniform<float>(rng, n, out, lo, hi): T = 128 * ceil(ceil(n/4)
    / 128) threads, thread t draws stream values 4t..4t+3 for elements t + j*T;
    value v -> v * (hi - lo) + lo with one rounding (the device contracts it to
    an FMA; in float64 the product and sum are exact here, so one float32
    rounding reproduces it).  The stream advances by n afterwards."""
    T = ((n + 3) // 4 + 127) // 128 * 128
    start = Pcg32.__new__(Pcg32)
    start.state, start.inc = rng.state, rng.inc
    u = start.next_uints(4 * T)
    v = (u >> np.uint32(9)).astype(np.float64) * 2.0 ** -23
    vals = v.reshape(T, 4).T.reshape(-1)[:n]
    lo32, hi32 = np.float32(lo), np.float32(hi)
    span = np.float64(np.float32(hi32 - lo32))
    out = (vals * span + np.float64(lo32)).astype(np.float32)
    rng.advance(n)
    return out


def tcnn_initial_params(layout: Layout, seed: int = 1337) -> np.ndarray:
    """The float32 parameter vector tcnn's Trainer(model, ..., seed) starts from."""
    rng = Pcg32(seed_seq_generate([seed], 2)[0])
    n = layout.net
    out = np.zeros(layout.n_params, np.float32)
    shapes = [(n.width, n.n_input)] + [(n.width, n.width)] * (n.n_hidden_layers - 1) + [(n.n_output_padded, n.width)]
    off = 0
    two = np.float32(2.0)
    for rows, cols in shapes:
        scale = np.sqrt(np.float32(6.0) / np.float32(rows + cols), dtype=np.float32)
        for k in range(rows * cols):
            out[off + k] = rng.next_float() * two * scale - scale
        off += rows * cols
    grids = {}
    for i, s in enumerate(layout.segs):
        if s.type == SEG_GRID and s.param_offset not in grids:
            offs, _, _ = layout.grid_levels(i)
            grids[s.param_offset] = int(offs[-1]) * s.n_features
    for p0 in sorted(grids):
        cnt = grids[p0]
        out[p0:p0 + cnt] = uniform_table(rng, cnt, -1e-4, 1e-4)
    return out
