"""voxelize(mesh, res): host side of the GPU voxelizer.

Mirrors the reference CLI's GPU branch (neural_lod/cuda_voxelizer/src/main.cpp:
199-258): per-vertex bounding box, cubic box with optional 1e-4 padding
(``-pad``), voxinfo, then the surface (default) or solid (``-solid``) kernel
writing a bit table with bit 31-(i%32) of word i/32 set for voxel
i = x + y*G + z*G*G.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import VoxInfo

SURFACE = 0
SOLID = 1 << 0
MORTON = 1 << 1
PAD_EPS = 1 << 2
ACCUMULATE = 1 << 3
DIRECT = 1 << 4  # one triangle per lane, global atomics
TILED = 1 << 5   # LDS-tiled whenever the table allows it
_PATHS = {"auto": 0, "direct": DIRECT, "tiled": TILED}


def _flags(solid: bool, pad: bool, morton: bool, accumulate: bool = False, path: str = "auto") -> int:
    if path not in _PATHS:
        raise ValueError(f"path must be one of {sorted(_PATHS)}")
    return (SOLID if solid else 0) | (PAD_EPS if pad else 0) | (MORTON if morton else 0) | (
        ACCUMULATE if accumulate else 0) | _PATHS[path]


def table_words(res: int) -> int:
    out = C.c_uint64()
    _lib.call("nlod_vox_table_words", res, C.byref(out))
    return out.value


def voxinfo_from_bbox(bbox_min, bbox_max, res: int, pad: bool = False, n_triangles: int = 0) -> VoxInfo:
    mn = (C.c_float * 3)(*[float(v) for v in bbox_min])
    mx = (C.c_float * 3)(*[float(v) for v in bbox_max])
    info = VoxInfo()
    _lib.call("nlod_voxinfo_from_bbox", mn, mx, res, _flags(False, pad, False), n_triangles, C.byref(info))
    return info


def _check_tris(triangles: torch.Tensor) -> torch.Tensor:
    if not (isinstance(triangles, torch.Tensor) and triangles.is_cuda):
        raise TypeError("triangles must be a CUDA (HIP) float32 tensor of shape [n, 9] or [n, 3, 3]")
    if triangles.dtype != torch.float32:
        raise TypeError("triangles must be float32")
    t = triangles.reshape(-1, 9)
    if not t.is_contiguous():
        t = t.contiguous()
    return t


def mesh_bbox(triangles: torch.Tensor) -> tuple[np.ndarray, np.ndarray]:
    t = _check_tris(triangles)
    mn = (C.c_float * 3)()
    mx = (C.c_float * 3)()
    stream = torch.cuda.current_stream(t.device).cuda_stream
    _lib.call("nlod_mesh_bbox", C.c_void_p(stream), C.c_void_p(t.data_ptr()), t.shape[0], mn, mx)
    return np.array(mn, np.float32), np.array(mx, np.float32)


def voxelize(triangles: torch.Tensor, res: int, solid: bool = False, pad: bool = False, morton: bool = False,
             info: VoxInfo | None = None, out: torch.Tensor | None = None,
             path: str = "auto") -> tuple[torch.Tensor, VoxInfo]:
    """Voxelize a device triangle soup; returns (int32 bit table on device, voxinfo).

    The table holds ceil(res^3/32) 32-bit words (reinterpret as uint32).
    With ``info`` given, that box is used as is (no bounding-box pass).
    ``path`` picks the kernels ("direct": one triangle per lane with global
    atomics; "tiled": LDS tiles of voxel rows; "auto": tiled for larger
    row-major jobs); all give identical bits."""
    t = _check_tris(triangles)
    words = table_words(res)
    if out is None:
        out = torch.empty(words, dtype=torch.int32, device=t.device)
    elif out.numel() < words or out.dtype != torch.int32 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous int32 tensor of at least {words} words")
    stream = C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)
    flags = _flags(solid, pad, morton, path=path)
    if info is None:
        info = VoxInfo()
        _lib.call("nlod_voxelize_mesh", stream, C.c_void_p(t.data_ptr()), t.shape[0], res, flags,
                  C.c_void_p(out.data_ptr()), C.byref(info))
    else:
        info.n_triangles = t.shape[0]
        _lib.call("nlod_voxelize", stream, C.byref(info), C.c_void_p(t.data_ptr()), C.c_void_p(out.data_ptr()),
                  flags)
    return out, info


def unpack(table: torch.Tensor | np.ndarray, res: int) -> np.ndarray:
    """bool[z][y][x] occupancy from a bit table (host copy)."""
    a = table.cpu().numpy() if isinstance(table, torch.Tensor) else np.asarray(table)
    bits = np.unpackbits(a.view(np.uint32).astype(">u4").view(np.uint8))[: res ** 3]
    return bits.reshape(res, res, res).astype(bool)
