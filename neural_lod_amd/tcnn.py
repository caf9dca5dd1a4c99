"""tcnn-style operator surface over libnlod (host side of the C ABI).

Mirrors the parts of tiny-cuda-nn the neural-LoD path uses:
``NetworkWithInputEncoding`` (include/tiny-cuda-nn/network_with_input_encoding.h)
with inference / forward / backward, and a ``torch.autograd.Function``
harness in the style of tinycudann's ``_module_function``
(bindings/torch/tinycudann/modules.py:59-128): gradients of the output are
scaled by ``loss_scale`` before the native backward and divided afterwards.

Device memory and streams come from torch; every computation runs in
libnlod's HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np
import torch

from . import _lib
from ._lib import AdamConfig, ModelInfo, TrainerStatus

FP32 = 0
FP16 = 1
GRAD_IGNORE, GRAD_OVERWRITE, GRAD_ACCUMULATE = 0, 1, 2


def _stream(device=None):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t: torch.Tensor | None):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _as_json(cfg) -> bytes:
    return (cfg if isinstance(cfg, str) else json.dumps(cfg)).encode()


def _check_f32(t: torch.Tensor, name: str, shape=None):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise TypeError(f"{name} must be a contiguous float32 CUDA tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")


class Context:
    """Saved forward state (owned by libnlod; freed with the object)."""

    def __init__(self, handle: C.c_void_p, model: "NetworkWithInputEncoding", n: int):
        self._h = handle
        self.model = model
        self.n = n

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                _lib.lib().nlod_ctx_destroy(self._h)
        except Exception:  # interpreter shutdown
            pass
        self._h = None


class NetworkWithInputEncoding:
    """Model = encoding (hash grid / SH / identity composites) + fully fused MLP.

    Parameters are one float32 vector (network matrices first, then encoding
    tables) bound with :meth:`set_params`; the model keeps a reference to the
    bound tensors."""

    def __init__(self, n_input_dims: int, n_output_dims: int, encoding_config, network_config,
                 precision: int = FP32):
        h = C.c_void_p()
        _lib.call("nlod_model_create", n_input_dims, n_output_dims, _as_json(encoding_config),
                  _as_json(network_config), precision, C.byref(h))
        self._h = h
        info = ModelInfo()
        _lib.call("nlod_model_get_info", h, C.byref(info))
        self.n_params = info.n_params
        self.n_network_params = info.n_network_params
        self.n_input_dims = info.input_width
        self.encoded_width = info.encoded_width
        self.padded_output_width = info.padded_output_width
        self.n_output_dims = info.output_width
        self.network_width = info.network_width
        self.n_hidden_layers = info.n_hidden_layers
        self.precision = info.precision
        self._params = None
        self._inference_params = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                _lib.lib().nlod_model_destroy(self._h)
        except Exception:  # interpreter shutdown
            pass
        self._h = None

    # -- layout / parameters ----------------------------------------------------
    def describe(self) -> dict:
        need = C.c_size_t()
        _lib.call("nlod_model_describe", self._h, None, 0, C.byref(need))
        buf = C.create_string_buffer(need.value)
        _lib.call("nlod_model_describe", self._h, buf, need.value, C.byref(need))
        return json.loads(buf.value.decode())

    def initial_params(self, seed: int = 1337) -> np.ndarray:
        """tcnn Trainer initialisation (host float32[n_params])."""
        out = np.zeros(self.n_params, np.float32)
        _lib.call("nlod_model_init_params", self._h, seed, out.ctypes.data_as(C.c_void_p))
        return out

    def set_params(self, params: torch.Tensor, inference_params: torch.Tensor | None = None) -> None:
        _check_f32(params, "params", (self.n_params,))
        if inference_params is not None:
            _check_f32(inference_params, "inference_params", (self.n_params,))
        _lib.call("nlod_model_set_params", self._h, _stream(params.device), _ptr(params), _ptr(inference_params))
        self._params = params
        self._inference_params = inference_params

    # -- compute ------------------------------------------------------------------
    def inference(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        _check_f32(x, "input")
        n = x.shape[0]
        if x.dim() != 2 or x.shape[1] != self.n_input_dims:
            raise ValueError(f"input must be [n, {self.n_input_dims}]")
        if out is None:
            out = torch.empty((n, self.n_output_dims), dtype=torch.float32, device=x.device)
        _check_f32(out, "output", (n, self.n_output_dims))
        _lib.call("nlod_model_inference", self._h, _stream(x.device), n, _ptr(x), _ptr(out))
        return out

    def encode_fragments(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """FP16 models: encoding pass alone into the MFMA fragment layout (uint8 buffer)."""
        _check_f32(x, "input")
        n = x.shape[0]
        nb = C.c_uint64()
        _lib.call("nlod_model_feature_bytes", self._h, n, C.byref(nb))
        if out is None:
            out = torch.empty(nb.value, dtype=torch.uint8, device=x.device)
        elif out.numel() * out.element_size() < nb.value:
            raise ValueError(f"feature buffer needs {nb.value} bytes")
        _lib.call("nlod_model_encode_fragments", self._h, _stream(x.device), n, _ptr(x), _ptr(out))
        return out

    def mlp_fragments(self, features: torch.Tensor, n: int, out: torch.Tensor | None = None) -> torch.Tensor:
        """FP16 models: the fused MLP alone on encode_fragments' features of n queries."""
        if out is None:
            out = torch.empty((n, self.n_output_dims), dtype=torch.float32, device=features.device)
        _check_f32(out, "output", (n, self.n_output_dims))
        _lib.call("nlod_model_mlp_fragments", self._h, _stream(features.device), n, _ptr(features), _ptr(out))
        return out

    def encode(self, x: torch.Tensor, use_inference_params: bool = True) -> torch.Tensor:
        """Encoded features as [encoded_width, n] (SoA)."""
        _check_f32(x, "input")
        n = x.shape[0]
        enc = torch.empty((self.encoded_width, n), dtype=torch.float32, device=x.device)
        _lib.call("nlod_model_encode", self._h, _stream(x.device), n, _ptr(x), _ptr(enc), int(use_inference_params))
        return enc

    def encode_backward(self, x: torch.Tensor, d_encoded: torch.Tensor, d_params: torch.Tensor | None = None,
                        mode: int = GRAD_OVERWRITE) -> torch.Tensor:
        _check_f32(x, "input")
        _check_f32(d_encoded, "d_encoded", (self.encoded_width, x.shape[0]))
        if d_params is None:
            d_params = torch.empty(self.n_params, dtype=torch.float32, device=x.device)
        _lib.call("nlod_model_encode_backward", self._h, _stream(x.device), x.shape[0], _ptr(x), _ptr(d_encoded),
                  _ptr(d_params), mode)
        return d_params

    def forward(self, x: torch.Tensor, use_inference_params: bool = False,
                prepare_input_gradients: bool = False) -> tuple[torch.Tensor, Context]:
        _check_f32(x, "input")
        n = x.shape[0]
        out = torch.empty((n, self.n_output_dims), dtype=torch.float32, device=x.device)
        h = C.c_void_p()
        _lib.call("nlod_model_forward", self._h, _stream(x.device), n, _ptr(x), _ptr(out), int(use_inference_params),
                  int(prepare_input_gradients), C.byref(h))
        return out, Context(h, self, n)

    def backward(self, ctx: Context, x: torch.Tensor, d_out: torch.Tensor, d_params: torch.Tensor | None = None,
                 input_gradients: bool = False, mode: int = GRAD_OVERWRITE):
        _check_f32(d_out, "d_out", (ctx.n, self.n_output_dims))
        if d_params is None and mode != GRAD_IGNORE:
            d_params = torch.empty(self.n_params, dtype=torch.float32, device=x.device)
        d_in = torch.empty_like(x) if input_gradients else None
        _lib.call("nlod_model_backward", self._h, _stream(x.device), ctx._h, _ptr(x), _ptr(d_out), _ptr(d_params),
                  _ptr(d_in), mode)
        return d_params, d_in


LOSS_L2, LOSS_RELATIVE_L2, LOSS_BCE = 0, 1, 2
TRAINER_GRAPH = 1
TRAINER_UNFUSED = 2


class Trainer:
    """tcnn ``Trainer`` (include/tiny-cuda-nn/trainer.h): FP32 master
    parameters, Adam (+ ExponentialDecay / Ema wrappers from the optimizer
    JSON), a loss from the loss JSON, and dynamic loss scaling.  The model's
    parameters are bound to the trainer's buffers while the trainer lives."""

    def __init__(self, model: NetworkWithInputEncoding, optimizer_config, loss_config, seed: int = 1337,
                 loss_scale_config=None, graphs: bool = False, fused: bool = True):
        h = C.c_void_p()
        flags = (TRAINER_GRAPH if graphs else 0) | (0 if fused else TRAINER_UNFUSED)
        _lib.call("nlod_trainer_create", model._h, _as_json(optimizer_config), _as_json(loss_config),
                  _as_json(loss_scale_config) if loss_scale_config is not None else None, seed, flags, C.byref(h))
        self._h = h
        self.model = model
        model._params = None
        model._inference_params = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                _lib.lib().nlod_trainer_destroy(self._h)
        except Exception:  # interpreter shutdown
            pass
        self._h = None

    def training_step(self, x: torch.Tensor, target: torch.Tensor) -> None:
        """One forward + loss + backward + optimizer step (asynchronous)."""
        _check_f32(x, "input", (x.shape[0], self.model.n_input_dims))
        _check_f32(target, "target", (x.shape[0], self.model.n_output_dims))
        _lib.call("nlod_trainer_step", self._h, _stream(x.device), x.shape[0], _ptr(x), _ptr(target))

    def status(self) -> dict:
        st = TrainerStatus()
        _lib.call("nlod_trainer_get_status", self._h, C.byref(st))
        return {k: getattr(st, k) for k, _ in TrainerStatus._fields_ if k != "reserved"}

    def loss(self) -> float:
        """Loss of the last step (synchronizes)."""
        return self.status()["loss"]

    def _copy(self, which: int) -> torch.Tensor:
        out = torch.empty(self.model.n_params, dtype=torch.float32, device="cuda")
        _lib.call("nlod_trainer_copy", self._h, _stream(out.device), which, _ptr(out))
        return out

    def params(self) -> torch.Tensor:
        return self._copy(0)

    def inference_params(self) -> torch.Tensor:
        return self._copy(1)

    def gradients(self) -> torch.Tensor:
        return self._copy(2)

    def moments(self) -> tuple[torch.Tensor, torch.Tensor]:
        return self._copy(3), self._copy(4)

    def set_params(self, params: torch.Tensor) -> None:
        _check_f32(params, "params", (self.model.n_params,))
        _lib.call("nlod_trainer_set_params", self._h, _stream(params.device), _ptr(params))

    def serialize(self, with_optimizer_state: bool = False) -> bytes:
        need = C.c_size_t()
        _lib.call("nlod_trainer_serialize", self._h, int(with_optimizer_state), None, 0, C.byref(need))
        buf = C.create_string_buffer(need.value)
        _lib.call("nlod_trainer_serialize", self._h, int(with_optimizer_state), buf, need.value, C.byref(need))
        return buf.raw

    def deserialize(self, blob: bytes) -> None:
        _lib.call("nlod_trainer_deserialize", self._h, C.c_char_p(blob), len(blob))

    def weights_json(self) -> str:
        """The reference's weight file text (tcnn Trainer::serialize as JSON)."""
        need = C.c_size_t()
        _lib.call("nlod_trainer_weights_json", self._h, None, 0, C.byref(need))
        buf = C.create_string_buffer(need.value)
        _lib.call("nlod_trainer_weights_json", self._h, buf, need.value, C.byref(need))
        return buf.value.decode()

    def load_weights_json(self, text: str | bytes) -> None:
        """Load a reference weight file (params_type float or __half)."""
        data = text.encode() if isinstance(text, str) else bytes(text)
        _lib.call("nlod_trainer_load_weights_json", self._h, _stream(torch.device("cuda")), data, len(data))


def loss_evaluate(loss_type: int, pred: torch.Tensor, target: torch.Tensor, loss_scale: float = 1.0):
    """Per-element loss values and scaled gradients (tcnn Loss::evaluate)."""
    _check_f32(pred, "pred")
    _check_f32(target, "target", tuple(pred.shape))
    values = torch.empty_like(pred)
    grads = torch.empty_like(pred)
    _lib.call("nlod_loss_evaluate", _stream(pred.device), loss_type, pred.numel(), loss_scale, _ptr(pred),
              _ptr(target), _ptr(values), _ptr(grads))
    return values, grads


def adam_step(config: dict, weights, gradients, m, v, steps, n_matrix: int, loss_scale: float,
              optimizer_step: int) -> None:
    """One Adam step in place on device tensors (tcnn adam_step kernel)."""
    cfg = AdamConfig(learning_rate=config.get("learning_rate", 1e-3), beta1=config.get("beta1", 0.9),
                     beta2=config.get("beta2", 0.999), epsilon=config.get("epsilon", 1e-8),
                     l2_reg=config.get("l2_reg", 1e-8), relative_decay=config.get("relative_decay", 0.0),
                     absolute_decay=config.get("absolute_decay", 0.0),
                     clipping_magnitude=config.get("clipping_magnitude", 0.0),
                     non_matrix_learning_rate_factor=config.get("non_matrix_learning_rate_factor", 1.0),
                     optimize_matrix_params=int(config.get("optimize_matrix_params", True)),
                     optimize_non_matrix_params=int(config.get("optimize_non_matrix_params", True)),
                     adabound=int(config.get("adabound", False)))
    for t, name in ((weights, "weights"), (gradients, "gradients"), (m, "m"), (v, "v")):
        _check_f32(t, name, (weights.numel(),))
    if not (steps.is_cuda and steps.dtype == torch.int32 and steps.numel() == weights.numel()):
        raise TypeError("steps must be an int32 CUDA tensor of the weights' size")
    _lib.call("nlod_adam_step", _stream(weights.device), C.byref(cfg), weights.numel(), n_matrix, loss_scale,
              optimizer_step, _ptr(weights), _ptr(gradients), _ptr(m), _ptr(v), _ptr(steps))


class _ModuleFunction(torch.autograd.Function):
    """Autograd bridge (tinycudann modules.py:59-128 pattern)."""

    @staticmethod
    def forward(ctx, model: NetworkWithInputEncoding, x, params, loss_scale):
        ctx.set_materialize_grads(False)
        model.set_params(params)
        out, native_ctx = model.forward(x)
        ctx.model = model
        ctx.native_ctx = native_ctx
        ctx.loss_scale = loss_scale
        ctx.save_for_backward(x, params)
        return out

    @staticmethod
    def backward(ctx, d_out):
        if d_out is None:
            return None, None, None, None
        x, params = ctx.saved_tensors
        model = ctx.model
        want_in = ctx.needs_input_grad[1]
        scaled = (d_out * ctx.loss_scale).contiguous()
        d_params, d_in = model.backward(ctx.native_ctx, x, scaled, input_gradients=want_in)
        d_params = d_params / ctx.loss_scale
        if d_in is not None:
            d_in = d_in / ctx.loss_scale
        return None, d_in, d_params, None


class Module(torch.nn.Module):
    """torch.nn.Module wrapper: parameters are one float32 vector."""

    def __init__(self, n_input_dims, n_output_dims, encoding_config, network_config, seed=1337,
                 precision=FP32, loss_scale=1.0):
        super().__init__()
        self.native = NetworkWithInputEncoding(n_input_dims, n_output_dims, encoding_config, network_config, precision)
        self.params = torch.nn.Parameter(torch.from_numpy(self.native.initial_params(seed)))
        self.loss_scale = loss_scale

    def forward(self, x):
        return _ModuleFunction.apply(self.native, x.contiguous(), self.params, self.loss_scale)


class Wavefront:
    """Per-wavefront visibility + throughput queries (libnlod's wavefront
    engine): the visibility network, then the throughput network where the
    visibility reaches ``threshold``."""

    def __init__(self, visibility: NetworkWithInputEncoding, throughput: NetworkWithInputEncoding,
                 threshold: float = 0.5):
        h = C.c_void_p()
        _lib.call("nlod_wavefront_create", visibility._h, throughput._h, threshold, C.byref(h))
        self._h = h
        self.visibility = visibility
        self.throughput = throughput

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                _lib.lib().nlod_wavefront_destroy(self._h)
        except Exception:  # interpreter shutdown
            pass
        self._h = None

    def tune(self, max_n: int, device=None) -> None:
        _lib.call("nlod_wavefront_tune", self._h, _stream(device), max_n)

    def choice(self, n: int) -> int:
        c = C.c_int()
        _lib.call("nlod_wavefront_choice", self._h, n, C.byref(c))
        return c.value

    def max_shape(self) -> int:
        c = C.c_int()
        _lib.call("nlod_wavefront_max_shape", self._h, C.byref(c))
        return c.value

    def _outputs(self, n, device):
        return (torch.empty((n, self.visibility.n_output_dims), dtype=torch.float32, device=device),
                torch.empty((n, self.throughput.n_output_dims), dtype=torch.float32, device=device))

    def run(self, vis_in: torch.Tensor, thr_in: torch.Tensor, shape: int | None = None, out=None,
            thresholds: torch.Tensor | None = None):
        """thresholds: optional float32 [n], each segment's voxel threshold
        (LodGrid.march(..., thresholds=True)); default the engine's value."""
        n = vis_in.shape[0]
        _check_f32(vis_in, "vis_in", (n, self.visibility.n_input_dims))
        _check_f32(thr_in, "thr_in", (n, self.throughput.n_input_dims))
        vo, to = out if out is not None else self._outputs(n, vis_in.device)
        if thresholds is not None:
            _check_f32(thresholds, "thresholds", (n,))
            _lib.call("nlod_wavefront_run_thresholds", self._h, _stream(vis_in.device), n, _ptr(vis_in), _ptr(thr_in),
                      _ptr(thresholds), _ptr(vo), _ptr(to), -1 if shape is None else int(shape))
        elif shape is None:
            _lib.call("nlod_wavefront_run", self._h, _stream(vis_in.device), n, _ptr(vis_in), _ptr(thr_in), _ptr(vo),
                      _ptr(to))
        else:
            _lib.call("nlod_wavefront_run_shape", self._h, _stream(vis_in.device), n, _ptr(vis_in), _ptr(thr_in),
                      _ptr(vo), _ptr(to), shape)
        return vo, to

    def run_batches(self, counts, vis_in: torch.Tensor, thr_in: torch.Tensor, out=None, graph: bool = False,
                    thresholds: torch.Tensor | None = None):
        n = vis_in.shape[0]
        if sum(counts) != n:
            raise ValueError("counts must sum to the number of rows")
        _check_f32(vis_in, "vis_in", (n, self.visibility.n_input_dims))
        _check_f32(thr_in, "thr_in", (n, self.throughput.n_input_dims))
        vo, to = out if out is not None else self._outputs(n, vis_in.device)
        c = (C.c_uint32 * len(counts))(*counts)
        if thresholds is not None:
            _check_f32(thresholds, "thresholds", (n,))
            _lib.call("nlod_wavefront_run_batches_thresholds", self._h, _stream(vis_in.device), len(counts), c,
                      _ptr(vis_in), _ptr(thr_in), _ptr(thresholds), _ptr(vo), _ptr(to), int(graph))
        else:
            _lib.call("nlod_wavefront_run_batches", self._h, _stream(vis_in.device), len(counts), c, _ptr(vis_in),
                      _ptr(thr_in), _ptr(vo), _ptr(to), int(graph))
        return vo, to
