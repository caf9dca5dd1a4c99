"""The neural LoD path tracer (libnlod ``nlod_tracer_*``): one sample per
pixel through a LoD level, visibility network per voxel segment, throughput
network per bounce -- the reference's
WavefrontNeuralThroughputVisibilityLod::voxel_tracing
(rptr/cuda/neural-lod/wavefront_neural_throughput_visibility_lod.cu:464-720)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

AUTO, REFERENCE, PERSISTENT = 0, 1, 2


class Camera(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3),
        ("dir_top_left", C.c_float * 3),
        ("du", C.c_float * 3),
        ("dv", C.c_float * 3),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
    ]


class TracerParams(C.Structure):
    _fields_ = [
        ("lod", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("background_color", C.c_float * 3),
        ("envmap_color", C.c_float * 3),
        ("inference_min_extent_dilation", C.c_float),
        ("voxel_extent_dilation_outward", C.c_float),
        ("apply_rr", C.c_uint32),
        ("rr_start_bounce", C.c_uint32),
        ("apply_visibility_rr", C.c_uint32),
        ("stochastic_threshold", C.c_uint32),
        ("use_hg", C.c_uint32),
        ("hg_g", C.c_float),
        ("throughput_log_learning", C.c_uint32),
        ("schedule", C.c_uint32),
    ]


def default_params(**overrides) -> TracerParams:
    """The reference's defaults (max_depth 3, white background and
    environment, dilation 1e-4, no roulette, uniform sphere directions),
    with keyword overrides."""
    p = TracerParams()
    _lib.call("nlod_tracer_default_params", C.byref(p))
    for k, v in overrides.items():
        if k in ("background_color", "envmap_color"):
            getattr(p, k)[:] = [float(c) for c in v]
        else:
            setattr(p, k, v)
    return p


def look_at(eye, target, up, fov_y_deg: float, width: int, height: int) -> Camera:
    """Pinhole camera in the reference's ViewParams form: dir_top_left points
    at the top-left corner of the image plane at distance 1, du / dv span it."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    h = np.tan(np.radians(fov_y_deg) / 2.0)
    w = h * width / height
    du = 2.0 * w * r
    dv = -2.0 * h * u
    tl = f - w * r + h * u
    cam = Camera()
    cam.position[:] = [float(c) for c in eye]
    cam.dir_top_left[:] = [float(c) for c in tl]
    cam.du[:] = [float(c) for c in du]
    cam.dv[:] = [float(c) for c in dv]
    cam.width, cam.height = int(width), int(height)
    return cam


class Tracer:
    """Path tracer over level ``params.lod`` of a :class:`LodGrid` with a
    visibility and a throughput :class:`NetworkWithInputEncoding` (both keep
    their bound inference parameters)."""

    def __init__(self, grid, visibility, throughput):
        h = C.c_void_p()
        _lib.call("nlod_tracer_create", grid._h, visibility._h, throughput._h, C.byref(h))
        self._h = h
        self.grid, self.visibility, self.throughput = grid, visibility, throughput

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                _lib.lib().nlod_tracer_destroy(self._h)
        except Exception:  # interpreter shutdown
            pass
        self._h = None

    def schedule(self, params: TracerParams) -> int:
        out = C.c_int()
        _lib.call("nlod_tracer_schedule", self._h, C.byref(params), C.byref(out))
        return out.value

    def stats(self) -> dict:
        """Network evaluations of the last persistent frame and their tile occupancy."""
        out = (C.c_uint64 * 5)()
        _lib.call("nlod_tracer_stats", self._h, C.c_void_p(torch.cuda.current_stream(self.grid.device).cuda_stream),
                  out)
        vt, vq, tt, tq, rounds = (int(v) for v in out)
        return {"visibility_tiles": vt, "visibility_queries": vq, "throughput_tiles": tt, "throughput_queries": tq,
                "wave_rounds": rounds,
                "tile_occupancy": (vq + tq) / (16.0 * (vt + tt)) if vt + tt else 0.0}

    def render(self, camera: Camera, params: TracerParams, sample_idx: int = 0, sample_offset: int = 0,
               out=None, with_stats: bool = True):
        """(illumination float32 [h*w, 3], visibility queries int32 [h*w],
        final bounce int32 [h*w]) on the grid's device; ``out`` reuses buffers."""
        n = camera.width * camera.height
        dev = self.grid.device
        if out is None:
            illum = torch.empty((n, 3), dtype=torch.float32, device=dev)
            inf = torch.empty(n, dtype=torch.int32, device=dev) if with_stats else None
            bnc = torch.empty(n, dtype=torch.int32, device=dev) if with_stats else None
        else:
            illum, inf, bnc = out
        stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        ptr = lambda t: C.c_void_p(t.data_ptr() if t is not None else 0)  # noqa: E731
        _lib.call("nlod_tracer_render", self._h, stream, C.byref(camera), C.byref(params), int(sample_idx),
                  int(sample_offset), ptr(illum), ptr(inf), ptr(bnc))
        return illum, inf, bnc
