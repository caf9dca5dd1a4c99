"""LoD occupancy pyramid and first-active-voxel ray marching (libnlod
``nlod_lod_grid_*``): the voxel traversal of the reference's wavefront
renderer (first_active_voxel_intersection, neural_lod_learning_common.cuh:
784-832; raymarch_impl, wavefront_neural_pt.cuh:89-141) over bit tables
instead of nanovdb grids."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import VoxInfo

NO_VOXEL = 0xFFFFFFFF


class LodGrid:
    """Level 0 = a voxelization (bit table + its voxinfo); level l + 1 = OR of
    2x2x2 blocks of level l.  The grid keeps its own copy of the tables."""

    def __init__(self, table: torch.Tensor, info: VoxInfo, n_lods: int = 1):
        if not (isinstance(table, torch.Tensor) and table.is_cuda and table.dtype == torch.int32):
            raise TypeError("table must be the int32 CUDA bit table of a voxelization")
        h = C.c_void_p()
        stream = C.c_void_p(torch.cuda.current_stream(table.device).cuda_stream)
        _lib.call("nlod_lod_grid_create", stream, C.byref(info), C.c_void_p(table.data_ptr()), int(n_lods), C.byref(h))
        self._h = h
        self.device = table.device
        self.n_lods = int(n_lods)
        self.levels = []
        for lod in range(self.n_lods):
            ptr, li = C.c_void_p(), VoxInfo()
            _lib.call("nlod_lod_grid_level", h, lod, C.byref(ptr), C.byref(li))
            self.levels.append(li)

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                _lib.lib().nlod_lod_grid_destroy(self._h)
        except Exception:  # interpreter shutdown
            pass
        self._h = None

    def table(self, lod: int) -> torch.Tensor:
        """A copy of level `lod`'s bit table (int32 words)."""
        ptr, li = C.c_void_p(), VoxInfo()
        _lib.call("nlod_lod_grid_level", self._h, lod, C.byref(ptr), C.byref(li))
        words = (li.gridsize[0] ** 3 + 31) // 32

        class _View:  # zero-copy view of the grid's device table
            __cuda_array_interface__ = {"shape": (words,), "typestr": "<i4", "data": (ptr.value, False),
                                        "version": 3}

        return torch.as_tensor(_View(), device=self.device).clone()

    def active_voxels(self, lod: int) -> int:
        """Occupied voxels of level `lod` (the length of its threshold file)."""
        c = C.c_uint64()
        _lib.call("nlod_lod_grid_active_voxels", self._h, C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream),
                  int(lod), C.byref(c))
        return int(c.value)

    def set_thresholds(self, lod: int, thresholds: torch.Tensor, min_threshold: float = 0.1,
                       max_threshold: float = 0.9) -> None:
        """Per-voxel thresholds of level `lod`, one per occupied voxel in
        increasing Morton order (the layout of the reference's
        *_lod_<l>_optimal_thresholds.bin files), clamped to [min, max]
        (defaults: threshold_min_value / threshold_max_value,
        rptr/cuda/neural-lod/neural_lod_learning.h:223-224)."""
        t = thresholds
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
                and t.dim() == 1):
            raise TypeError("thresholds must be a contiguous 1-D float32 CUDA tensor")
        _lib.call("nlod_lod_grid_set_thresholds", self._h, C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream),
                  int(lod), t.numel(), C.c_void_p(t.data_ptr()), float(min_threshold), float(max_threshold))

    def march(self, lod: int, origins: torch.Tensor, dirs: torch.Tensor, t_min: torch.Tensor | None = None,
              prev: torch.Tensor | None = None, thresholds: bool = False):
        """First occupied voxel along each ray: (voxel int64 [n] with -1 for
        none, t_near [n], t_far [n]), plus the voxel's threshold [n] (-1 for
        none) when `thresholds`."""
        n = origins.shape[0]
        for t, name in ((origins, "origins"), (dirs, "dirs")):
            if t.shape != (n, 3) or t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
                raise ValueError(f"{name} must be a contiguous float32 CUDA tensor [n, 3]")
        vox = torch.empty(n, dtype=torch.int32, device=origins.device)
        tn = torch.empty(n, dtype=torch.float32, device=origins.device)
        tf = torch.empty(n, dtype=torch.float32, device=origins.device)
        stream = C.c_void_p(torch.cuda.current_stream(origins.device).cuda_stream)
        args = (self._h, stream, int(lod), n, C.c_void_p(origins.data_ptr()), C.c_void_p(dirs.data_ptr()),
                C.c_void_p(t_min.data_ptr()) if t_min is not None else None,
                C.c_void_p(prev.data_ptr()) if prev is not None else None, C.c_void_p(vox.data_ptr()),
                C.c_void_p(tn.data_ptr()), C.c_void_p(tf.data_ptr()))
        thr = None
        if thresholds:
            thr = torch.empty(n, dtype=torch.float32, device=origins.device)
            _lib.call("nlod_lod_grid_march_thresholds", *args, C.c_void_p(thr.data_ptr()))
        else:
            _lib.call("nlod_lod_grid_march", *args)
        v = vox.to(torch.int64) & 0xFFFFFFFF
        v[v == NO_VOXEL] = -1
        return (v, tn, tf, thr) if thresholds else (v, tn, tf)

