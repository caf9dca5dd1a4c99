"""ORACLE — test infrastructure only.

Python access to the CPU restatement of the reference's wavefront path-tracer
loop (``tracer_oracle.c``: voxel_tracing and its stage kernels,
rptr/cuda/neural-lod/wavefront_neural_throughput_visibility_lod.cu:464-720).
The networks are a Python callback ``net(which, inputs) -> outputs`` (which 0:
visibility [n, 6] -> [n, >= 1]; 1: throughput [n, 9 or 10] -> [n, >= 3]).
Only tests/ import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import oracle as O

NET_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float))


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("dir_top_left", C.c_float * 3), ("du", C.c_float * 3),
                ("dv", C.c_float * 3), ("width", C.c_uint32), ("height", C.c_uint32)]


class Params(C.Structure):
    _fields_ = [
        ("lod", C.c_uint32), ("max_depth", C.c_uint32),
        ("background", C.c_float * 3), ("envmap", C.c_float * 3),
        ("inference_min_extent_dilation", C.c_float), ("voxel_extent_dilation_outward", C.c_float),
        ("apply_rr", C.c_uint32), ("rr_start_bounce", C.c_uint32), ("apply_visibility_rr", C.c_uint32),
        ("stochastic_threshold", C.c_uint32), ("use_hg", C.c_uint32), ("hg_g", C.c_float),
        ("log_learning", C.c_uint32), ("n_thr_inputs", C.c_uint32), ("grid_res", C.c_uint32),
    ]


class Scene(C.Structure):
    _fields_ = [("bits", C.c_void_p), ("thr", C.c_void_p), ("R", C.c_uint32), ("bmin", C.c_float * 3),
                ("bmax", C.c_float * 3), ("voxel_size", C.c_float)]


class MarchRay(C.Structure):
    """nlo_ray (the oracle's per-path state), for the march hook."""
    _fields_ = [("o", C.c_float * 3), ("d", C.c_float * 3), ("rng", C.c_uint32), ("state", C.c_int),
                ("vox", C.c_int * 3), ("t_near", C.c_float), ("t_far", C.c_float), ("thr", C.c_float),
                ("T", C.c_float * 3), ("wi", C.c_float * 3), ("pdf_inv", C.c_float), ("illum", C.c_float * 3),
                ("pixel", C.c_uint32), ("needed", C.c_uint32), ("bounce", C.c_uint32)]


def _lib():
    L = O.lib()
    if not getattr(L, "_tracer_ready", False):
        L.nlo_trace.argtypes = [C.POINTER(Scene), C.POINTER(Camera), C.POINTER(Params), C.c_uint32, C.c_uint32,
                                NET_FN, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.nlo_lcg_seed.restype = C.c_uint32
        L.nlo_lcg_seed.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.nlo_lcg_next.restype = C.c_float
        L.nlo_lcg_next.argtypes = [C.POINTER(C.c_uint32)]
        L.nlo_sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.nlo_exp.restype = C.c_double
        L.nlo_exp.argtypes = [C.c_double]
        L.nlo_tracer_sample_dir.argtypes = [C.c_float, C.c_float, C.c_int, C.POINTER(C.c_float), C.c_float,
                                            C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.nlo_tracer_march.restype = C.c_int
        L.nlo_tracer_march.argtypes = [C.POINTER(Scene), C.POINTER(MarchRay)]
        L._tracer_ready = True
    return L


def scene(bits: np.ndarray, R: int, bmin, bmax, voxel_size: float, thr: np.ndarray | None = None) -> Scene:
    s = Scene()
    s._bits = np.ascontiguousarray(bits).view(np.uint32)
    s.bits = s._bits.ctypes.data
    s._thr = None if thr is None else np.ascontiguousarray(thr, np.float32)
    s.thr = 0 if thr is None else s._thr.ctypes.data
    s.R = R
    s.bmin[:] = [float(v) for v in bmin]
    s.bmax[:] = [float(v) for v in bmax]
    s.voxel_size = float(voxel_size)
    return s


def params_from(p, n_thr_inputs: int, grid_res: int) -> Params:
    """Oracle parameters from the library's TracerParams structure."""
    q = Params()
    q.lod, q.max_depth = p.lod, p.max_depth
    q.background[:] = list(p.background_color)
    q.envmap[:] = list(p.envmap_color)
    q.inference_min_extent_dilation = p.inference_min_extent_dilation
    q.voxel_extent_dilation_outward = p.voxel_extent_dilation_outward
    q.apply_rr, q.rr_start_bounce = p.apply_rr, p.rr_start_bounce
    q.apply_visibility_rr, q.stochastic_threshold = p.apply_visibility_rr, p.stochastic_threshold
    q.use_hg, q.hg_g, q.log_learning = p.use_hg, p.hg_g, p.throughput_log_learning
    q.n_thr_inputs, q.grid_res = n_thr_inputs, grid_res
    return q


def camera_from(cam) -> Camera:
    c = Camera()
    for f in ("position", "dir_top_left", "du", "dv"):
        getattr(c, f)[:] = list(getattr(cam, f))
    c.width, c.height = cam.width, cam.height
    return c


def trace(sc: Scene, cam: Camera, p: Params, net, sample_idx: int = 0, sample_offset: int = 0):
    """(illumination [n, 3] float32, inferences [n] uint32, bounces [n] uint32)."""
    n = cam.width * cam.height
    n_in = {0: 6, 1: p.n_thr_inputs}
    n_out = {0: 1, 1: 3}

    def cb(_user, which, count, inp, out):
        x = np.ctypeslib.as_array(inp, shape=(count * n_in[which],)).reshape(count, n_in[which]).copy()
        y = np.asarray(net(which, x), np.float32)[:, : n_out[which]]
        dst = np.ctypeslib.as_array(out, shape=(count * n_out[which],))
        dst[:] = np.ascontiguousarray(y).reshape(-1)

    fn = NET_FN(cb)
    illum = np.full((n, 3), np.nan, np.float32)
    inf = np.zeros(n, np.uint32)
    bnc = np.zeros(n, np.uint32)
    _lib().nlo_trace(C.byref(sc), C.byref(cam), C.byref(p), sample_idx, sample_offset, fn, None,
                     illum.ctypes.data, inf.ctypes.data, bnc.ctypes.data)
    return illum, inf, bnc


def lcg_seed(index: int, frame: int, linear: int) -> int:
    return int(_lib().nlo_lcg_seed(index, frame, linear))


def lcg_stream(state: int, count: int) -> tuple[list[int], list[float]]:
    """States after each step and the floats drawn."""
    s = C.c_uint32(state)
    states, floats = [], []
    for _ in range(count):
        floats.append(float(_lib().nlo_lcg_next(C.byref(s))))
        states.append(int(s.value))
    return states, floats


def sincos(x: float) -> tuple[float, float]:
    s, c = C.c_double(), C.c_double()
    _lib().nlo_sincos(x, C.byref(s), C.byref(c))
    return s.value, c.value


def exp(x: float) -> float:
    return float(_lib().nlo_exp(x))


def sample_dir(u0: float, u1: float, hg: bool = False, w_o=(0.0, 0.0, 1.0), g: float = 0.5):
    """(direction float32[3], pdf_inv) of the throughput query's wi sampler."""
    wo = (C.c_float * 3)(*w_o)
    out = (C.c_float * 3)()
    pdf_inv = C.c_float()
    _lib().nlo_tracer_sample_dir(u0, u1, int(hg), wo, g, out, C.byref(pdf_inv))
    return np.array(out, np.float32), float(pdf_inv.value)


def march(sc: Scene, origin, direction, t_near=-1.0, t_far=-1.0, prev=(-1, -1, -1)):
    """One raymarch_impl step in index space: (hit, voxel (x, y, z), t_near, t_far)."""
    r = MarchRay()
    r.o[:] = [float(v) for v in origin]
    r.d[:] = [float(v) for v in direction]
    r.t_near, r.t_far = float(t_near), float(t_far)
    r.vox[:] = [int(v) for v in prev]
    hit = _lib().nlo_tracer_march(C.byref(sc), C.byref(r))
    return bool(hit), tuple(r.vox), float(r.t_near), float(r.t_far)
