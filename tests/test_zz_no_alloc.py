"""nlod_voxelize_noalloc against nlod_voxelize, through the C ABI.

The allocation-free entry point runs the same kernels with the tiled path's
scratch in a fixed device array; its tables must equal nlod_voxelize's bit for
bit, on consecutive calls whose second input is smaller (the array keeps the
previous call's contents), in both modes and on both kernel paths.  A job
whose scratch would not fit is rejected before anything runs."""
import ctypes as C

import numpy as np
import pytest
import torch

from neural_lod_amd import _lib, synthetic
from neural_lod_amd import voxelizer as V

pytestmark = pytest.mark.gpu


def _call(name, info, tris, table, flags):
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = _lib.lib()
    st = getattr(L, name)(stream, C.byref(info), C.c_void_p(tris.data_ptr()), C.c_void_p(table.data_ptr()), flags)
    return st


def _pair(tris_np, res, flags):
    tris = torch.from_numpy(tris_np).cuda()
    mn, mx = V.mesh_bbox(tris)
    info = V.voxinfo_from_bbox(mn, mx, res, n_triangles=tris.shape[0])
    words = V.table_words(res)
    ref = torch.full((words,), -1, dtype=torch.int32, device="cuda")
    got = torch.full((words,), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    assert _call("nlod_voxelize", info, tris, ref, flags) == 0, _lib.lib().nlod_last_error()
    assert _call("nlod_voxelize_noalloc", info, tris, got, flags) == 0, _lib.lib().nlod_last_error()
    torch.cuda.synchronize()
    return ref, got


@pytest.mark.parametrize("mode", [V.SOLID, V.SURFACE])
def test_noalloc_equals_voxelize_twice_shrinking(cuda, mode):
    # first a large job (tiled scratch of about 11 MiB), then a smaller one on
    # the same array; each must equal nlod_voxelize exactly
    big = synthetic.blob_mesh(200_000, seed=21)
    small = synthetic.blob_mesh(9_000, seed=22)
    for tris, res, path in ((big, 256, V.TILED), (small, 128, V.TILED), (small, 64, V.DIRECT)):
        ref, got = _pair(tris, res, mode | path)
        assert torch.equal(ref, got), f"noalloc differs (n={tris.shape[0]}, res={res}, flags={mode | path})"
        assert int((ref != 0).sum().item()) > 0


def test_noalloc_default_path_and_rejection(cuda):
    tris = torch.from_numpy(synthetic.blob_mesh(50_000, seed=23)).cuda()
    mn, mx = V.mesh_bbox(tris)
    info = V.voxinfo_from_bbox(mn, mx, 256, n_triangles=tris.shape[0])
    words = V.table_words(256)
    ref = torch.empty(words, dtype=torch.int32, device="cuda")
    got = torch.empty(words, dtype=torch.int32, device="cuda")
    assert _call("nlod_voxelize", info, tris, ref, V.SOLID) == 0
    assert _call("nlod_voxelize_noalloc", info, tris, got, V.SOLID) == 0
    torch.cuda.synchronize()
    assert torch.equal(ref, got)
    # 2M triangles at 1024^3 on the tiled path need far more than 32 MiB of scratch
    n_big = 2_000_000
    big_info = V.voxinfo_from_bbox(mn, mx, 1024, n_triangles=n_big)
    table = torch.full((V.table_words(1024),), 7, dtype=torch.int32, device="cuda")
    huge = torch.zeros((n_big, 9), dtype=torch.float32, device="cuda")
    st = _call("nlod_voxelize_noalloc", big_info, huge, table, V.SURFACE | V.TILED)
    assert st == 1, st  # NLOD_ERR_INVALID_ARGUMENT
    assert b"scratch" in _lib.lib().nlod_last_error()
    torch.cuda.synchronize()
    assert bool((table == 7).all().item()), "a rejected call must not touch the table"
