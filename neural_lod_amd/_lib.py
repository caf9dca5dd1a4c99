"""ctypes binding of libnlod.so (the C ABI declared in include/nlod.h).

The library must have been built (``python -m neural_lod_amd.build`` or
``__graft_entry__.build()``); there is no fallback implementation.  torch is
imported first so that libnlod shares torch's HIP runtime instance.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import torch  # noqa: F401  (loads the HIP runtime libnlod links against)

LIB_PATH = Path(__file__).resolve().parent / "libnlod.so"

_lock = threading.Lock()
_lib: C.CDLL | None = None

NLOD_OK = 0
STATUS_NAMES = {
    1: "NLOD_ERR_INVALID_ARGUMENT",
    2: "NLOD_ERR_CONFIG",
    3: "NLOD_ERR_UNSUPPORTED",
    4: "NLOD_ERR_HIP",
    5: "NLOD_ERR_OUT_OF_MEMORY",
    6: "NLOD_ERR_STATE",
}


class NlodError(RuntimeError):
    """A libnlod call returned a non-zero status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class VoxInfo(C.Structure):
    _fields_ = [
        ("bbox_min", C.c_float * 3),
        ("bbox_max", C.c_float * 3),
        ("gridsize", C.c_uint32 * 3),
        ("reserved0", C.c_uint32),
        ("n_triangles", C.c_uint64),
        ("unit", C.c_float * 3),
        ("reserved1", C.c_uint32),
    ]


_vp = C.c_void_p
_fp = C.POINTER(C.c_float)
_u32 = C.c_uint32
_u64 = C.c_uint64
_i32 = C.c_int

# name -> argtypes (every function returns int status unless listed in _RESTYPES)
SIGNATURES: dict[str, list] = {
    "nlod_last_error": [],
    "nlod_version": [],
    "nlod_vox_table_words": [_u32, C.POINTER(_u64)],
    "nlod_voxinfo_from_bbox": [_fp, _fp, _u32, _u32, _u64, C.POINTER(VoxInfo)],
    "nlod_mesh_bbox": [_vp, _vp, _u64, _fp, _fp],
    "nlod_voxelize": [_vp, C.POINTER(VoxInfo), _vp, _vp, _u32],
    "nlod_voxelize_mesh": [_vp, _vp, _u64, _u32, _u32, _vp, C.POINTER(VoxInfo)],
    "nlod_voxelize_noalloc": [_vp, C.POINTER(VoxInfo), _vp, _vp, _u32],
    "nlod_model_create": [_u32, _u32, C.c_char_p, C.c_char_p, _i32, C.POINTER(_vp)],
    "nlod_model_destroy": [_vp],
    "nlod_model_get_info": [_vp, _vp],
    "nlod_model_describe": [_vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)],
    "nlod_model_init_params": [_vp, _u32, _vp],
    "nlod_model_set_params": [_vp, _vp, _vp, _vp],
    "nlod_model_inference": [_vp, _vp, _u32, _vp, _vp],
    "nlod_model_forward": [_vp, _vp, _u32, _vp, _vp, _i32, _i32, C.POINTER(_vp)],
    "nlod_model_backward": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32],
    "nlod_ctx_destroy": [_vp],
    "nlod_model_encode": [_vp, _vp, _u32, _vp, _vp, _i32],
    "nlod_model_encode_backward": [_vp, _vp, _u32, _vp, _vp, _vp, _i32],
    "nlod_model_feature_bytes": [_vp, _u32, C.POINTER(_u64)],
    "nlod_model_encode_fragments": [_vp, _vp, _u32, _vp, _vp],
    "nlod_model_mlp_fragments": [_vp, _vp, _u32, _vp, _vp],
    "nlod_trainer_create": [_vp, C.c_char_p, C.c_char_p, C.c_char_p, _u32, _u32, C.POINTER(_vp)],
    "nlod_trainer_destroy": [_vp],
    "nlod_trainer_step": [_vp, _vp, _u32, _vp, _vp],
    "nlod_trainer_get_status": [_vp, _vp],
    "nlod_trainer_params": [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)],
    "nlod_trainer_copy": [_vp, _vp, _i32, _vp],
    "nlod_trainer_set_params": [_vp, _vp, _vp],
    "nlod_trainer_serialize": [_vp, _i32, _vp, C.c_size_t, C.POINTER(C.c_size_t)],
    "nlod_trainer_deserialize": [_vp, _vp, C.c_size_t],
    "nlod_trainer_weights_json": [_vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)],
    "nlod_trainer_load_weights_json": [_vp, _vp, C.c_char_p, C.c_size_t],
    "nlod_loss_evaluate": [_vp, _i32, _u64, C.c_float, _vp, _vp, _vp, _vp],
    "nlod_adam_step": [_vp, _vp, _u64, _u64, C.c_float, _u32, _vp, _vp, _vp, _vp, _vp],
    "nlod_wavefront_create": [_vp, _vp, C.c_float, C.POINTER(_vp)],
    "nlod_wavefront_destroy": [_vp],
    "nlod_wavefront_tune": [_vp, _vp, _u32],
    "nlod_wavefront_choice": [_vp, _u32, C.POINTER(_i32)],
    "nlod_wavefront_max_shape": [_vp, C.POINTER(_i32)],
    "nlod_wavefront_run": [_vp, _vp, _u32, _vp, _vp, _vp, _vp],
    "nlod_wavefront_run_shape": [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _i32],
    "nlod_wavefront_run_thresholds": [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _i32],
    "nlod_wavefront_run_batches": [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _i32],
    "nlod_wavefront_run_batches_thresholds": [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _i32],
    "nlod_labeler_create": [_vp, _vp, _u64, _u32, _u32, C.POINTER(_vp)],
    "nlod_labeler_destroy": [_vp],
    "nlod_labeler_info": [_vp, _u32, C.POINTER(VoxInfo), C.POINTER(_u32), C.POINTER(_u64)],
    "nlod_labeler_visibility": [_vp, _vp, _u64, _u64, _u32, _vp, _vp],
    "nlod_labeler_throughput": [_vp, _vp, _u64, _u64, _u32, _u32, _u32, _u64, _vp, _vp, _vp],
    "nlod_lod_grid_create": [_vp, C.POINTER(VoxInfo), _vp, _u32, C.POINTER(_vp)],
    "nlod_lod_grid_destroy": [_vp],
    "nlod_lod_grid_level": [_vp, _u32, C.POINTER(_vp), C.POINTER(VoxInfo)],
    "nlod_lod_grid_march": [_vp, _vp, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "nlod_lod_grid_march_thresholds": [_vp, _vp, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "nlod_lod_grid_active_voxels": [_vp, _vp, _u32, C.POINTER(_u64)],
    "nlod_lod_grid_set_thresholds": [_vp, _vp, _u32, _u64, _vp, C.c_float, C.c_float],
    "nlod_optimal_thresholds": [_vp, _u32, _u32, _vp, _vp, _u32, _i32, C.c_float, _vp, _vp],
    "nlod_tracer_default_params": [_vp],
    "nlod_tracer_create": [_vp, _vp, _vp, C.POINTER(_vp)],
    "nlod_tracer_destroy": [_vp],
    "nlod_tracer_render": [_vp, _vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp],
    "nlod_tracer_schedule": [_vp, _vp, C.POINTER(_i32)],
    "nlod_tracer_stats": [_vp, _vp, _vp],
}


class ModelInfo(C.Structure):
    _fields_ = [
        ("n_params", C.c_uint64),
        ("n_network_params", C.c_uint64),
        ("input_width", C.c_uint32),
        ("encoded_width", C.c_uint32),
        ("padded_output_width", C.c_uint32),
        ("output_width", C.c_uint32),
        ("network_width", C.c_uint32),
        ("n_hidden_layers", C.c_uint32),
        ("precision", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class TrainerStatus(C.Structure):
    _fields_ = [
        ("loss", C.c_double),
        ("grad_norm", C.c_double),
        ("loss_scale", C.c_float),
        ("learning_rate", C.c_float),
        ("step", C.c_uint32),
        ("skipped_steps", C.c_uint32),
        ("last_step_skipped", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class AdamConfig(C.Structure):
    _fields_ = [(k, C.c_float) for k in ("learning_rate", "beta1", "beta2", "epsilon", "l2_reg", "relative_decay",
                                          "absolute_decay", "clipping_magnitude",
                                          "non_matrix_learning_rate_factor")] + \
               [(k, C.c_int) for k in ("optimize_matrix_params", "optimize_non_matrix_params", "adabound")]


_RESTYPES = {"nlod_last_error": C.c_char_p, "nlod_version": C.c_char_p}


def lib() -> C.CDLL:
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"libnlod.so not found at {LIB_PATH}: build it first "
                    "(python -m neural_lod_amd.build); there is no CPU fallback")
            L = C.CDLL(str(LIB_PATH))
            for name, args in SIGNATURES.items():
                fn = getattr(L, name)
                fn.argtypes = args
                fn.restype = _RESTYPES.get(name, C.c_int)
            _lib = L
    return _lib


def call(name: str, *args) -> None:
    """Call a status-returning libnlod function; raise NlodError on failure."""
    L = lib()
    status = getattr(L, name)(*args)
    if status != NLOD_OK:
        msg = L.nlod_last_error().decode(errors="replace")
        raise NlodError(status, msg)


def exported_symbols() -> list[str]:
    return list(SIGNATURES)
