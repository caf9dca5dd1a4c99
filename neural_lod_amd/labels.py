"""Training samples from a mesh for the LoD networks (libnlod ``nlod_labeler_*``).

Visibility: rays start uniformly inside a uniformly chosen occupied voxel of
the level-0 grid with a direction uniform on the sphere; the label is the
renderer's visibility query (rptr/vulkan/raygen.rgen:690-730): 1 when the
nearest hit lies inside the voxel.  Throughput: the renderer's voxel-local
throughput query (raygen.rgen:505-688) averaged over spp paths, at a uniformly
chosen level of the LoD pyramid.  Inputs are the networks' rows as the
reference's training loop feeds them (neural_lod_learning.cu:588-598, 686-694).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import VoxInfo


class Labeler:
    """LoD pyramid over a mesh, its occupied voxels and their triangle lists.
    Keeps a reference to the triangle tensor, which the GPU kernels read."""

    def __init__(self, triangles: torch.Tensor, res: int = 64, n_lods: int = 1):
        if not (isinstance(triangles, torch.Tensor) and triangles.is_cuda and triangles.dtype == torch.float32):
            raise TypeError("triangles must be a CUDA float32 tensor of shape [n, 9] or [n, 3, 3]")
        self._tris = triangles.reshape(-1, 9).contiguous()
        self.res = int(res)
        self.n_lods = int(n_lods)
        h = C.c_void_p()
        stream = C.c_void_p(torch.cuda.current_stream(self._tris.device).cuda_stream)
        _lib.call("nlod_labeler_create", stream, C.c_void_p(self._tris.data_ptr()), self._tris.shape[0], self.res,
                  self.n_lods, C.byref(h))
        self._h = h
        self.levels = []
        for lod in range(self.n_lods):
            info, n_occ, n_list = VoxInfo(), C.c_uint32(), C.c_uint64()
            _lib.call("nlod_labeler_info", h, lod, C.byref(info), C.byref(n_occ), C.byref(n_list))
            self.levels.append({"grid": info, "n_occupied": n_occ.value, "n_list_entries": n_list.value})
        self.grid = self.levels[0]["grid"]
        self.n_occupied = self.levels[0]["n_occupied"]

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                _lib.lib().nlod_labeler_destroy(self._h)
        except Exception:  # interpreter shutdown
            pass
        self._h = None

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self._tris.device).cuda_stream)

    @staticmethod
    def _check(t: torch.Tensor, shape):
        if tuple(t.shape) != shape or t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError(f"expected a contiguous float32 tensor of shape {shape}")

    def sample(self, n: int, seed: int, first: int = 0, out=None) -> tuple[torch.Tensor, torch.Tensor]:
        """Visibility samples first .. first + n - 1: (inputs [n, 6], labels [n, 1])."""
        dev = self._tris.device
        if out is None:
            out = (torch.empty((n, 6), dtype=torch.float32, device=dev),
                   torch.empty((n, 1), dtype=torch.float32, device=dev))
        x, y = out
        self._check(x, (n, 6))
        self._check(y, (n, 1))
        _lib.call("nlod_labeler_visibility", self._h, self._stream(), int(seed) & (2 ** 64 - 1), int(first), int(n),
                  C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()))
        return x, y

    def throughput(self, n: int, seed: int, first: int = 0, spp: int = 64, max_depth: int = 3,
                   albedo_seed: int = 1, out=None) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """Throughput samples: (inputs [n, 10], RGB targets [n, 3], valid [n])."""
        dev = self._tris.device
        if out is None:
            out = (torch.empty((n, 10), dtype=torch.float32, device=dev),
                   torch.empty((n, 3), dtype=torch.float32, device=dev),
                   torch.empty((n,), dtype=torch.float32, device=dev))
        x, t, v = out
        self._check(x, (n, 10))
        self._check(t, (n, 3))
        self._check(v, (n,))
        _lib.call("nlod_labeler_throughput", self._h, self._stream(), int(seed) & (2 ** 64 - 1), int(first), int(n),
                  int(spp), int(max_depth), int(albedo_seed) & (2 ** 64 - 1), C.c_void_p(x.data_ptr()),
                  C.c_void_p(t.data_ptr()), C.c_void_p(v.data_ptr()))
        return x, t, v


# the visibility-only name used by callers that train the visibility network
VisibilityLabeler = Labeler
