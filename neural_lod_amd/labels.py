"""Visibility training samples from a mesh (libnlod ``nlod_labeler_*``).

Rays start uniformly inside a uniformly chosen occupied voxel of a res^3 grid
over the mesh (surface voxelization) with a direction uniform on the sphere;
the label is the renderer's visibility query (rptr/vulkan/raygen.rgen:690-730):
1 when the nearest hit lies inside the voxel.  Inputs are the network's
[position in the grid box, (direction + 1) / 2] rows, as the reference's
training loop feeds them (neural_lod_learning.cu:588-598).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import VoxInfo


class VisibilityLabeler:
    """Holds the grid, the occupied voxels and their triangle lists; keeps a
    reference to the triangle tensor, which the GPU kernel reads."""

    def __init__(self, triangles: torch.Tensor, res: int = 64):
        if not (isinstance(triangles, torch.Tensor) and triangles.is_cuda and triangles.dtype == torch.float32):
            raise TypeError("triangles must be a CUDA float32 tensor of shape [n, 9] or [n, 3, 3]")
        self._tris = triangles.reshape(-1, 9).contiguous()
        self.res = int(res)
        h = C.c_void_p()
        stream = C.c_void_p(torch.cuda.current_stream(self._tris.device).cuda_stream)
        _lib.call("nlod_labeler_create", stream, C.c_void_p(self._tris.data_ptr()), self._tris.shape[0], self.res,
                  C.byref(h))
        self._h = h
        info, n_occ, n_list = VoxInfo(), C.c_uint32(), C.c_uint64()
        _lib.call("nlod_labeler_info", h, C.byref(info), C.byref(n_occ), C.byref(n_list))
        self.grid = info
        self.n_occupied = n_occ.value
        self.n_list_entries = n_list.value

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                _lib.lib().nlod_labeler_destroy(self._h)
        except Exception:  # interpreter shutdown
            pass
        self._h = None

    def sample(self, n: int, seed: int, first: int = 0, out: tuple[torch.Tensor, torch.Tensor] | None = None
               ) -> tuple[torch.Tensor, torch.Tensor]:
        """Samples first .. first + n - 1 of the stream `seed`: (inputs [n, 6], labels [n, 1])."""
        dev = self._tris.device
        if out is None:
            out = (torch.empty((n, 6), dtype=torch.float32, device=dev),
                   torch.empty((n, 1), dtype=torch.float32, device=dev))
        x, y = out
        if x.shape != (n, 6) or y.shape != (n, 1) or not (x.is_contiguous() and y.is_contiguous()):
            raise ValueError("out must be contiguous float32 tensors of shapes [n, 6] and [n, 1]")
        stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.call("nlod_labeler_visibility", self._h, stream, int(seed) & (2 ** 64 - 1), int(first), int(n),
                  C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()))
        return x, y
