"""ORACLE — test infrastructure only.

Python access to the CPU oracle (``liboracle.so``, built from the C sources in
this directory by ``make -C oracle``).  Only tests/, ``__graft_entry__.smoke()``
and bench.py's CPU-baseline leg import this module, and only as the checker.
The product (``neural_lod_amd``) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None


class VoxInfo(C.Structure):
    _fields_ = [
        ("bbox_min", C.c_float * 3),
        ("bbox_max", C.c_float * 3),
        ("gridsize", C.c_uint32 * 3),
        ("pad0", C.c_uint32),
        ("n_triangles", C.c_uint64),
        ("unit", C.c_float * 3),
        ("pad1", C.c_uint32),
    ]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "-B"], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        fp = C.POINTER(C.c_float)
        u32p = C.POINTER(C.c_uint32)
        L.nlo_mesh_bbox.argtypes = [fp, C.c_uint64, fp, fp]
        L.nlo_voxinfo_from_bbox.argtypes = [fp, fp, C.c_uint32, C.c_float, C.c_uint64, C.POINTER(VoxInfo)]
        L.nlo_voxelize_surface.argtypes = [C.POINTER(VoxInfo), fp, u32p, C.c_int]
        L.nlo_voxelize_solid.argtypes = [C.POINTER(VoxInfo), fp, u32p, C.c_int]
        L.nlo_visibility_samples.restype = C.c_int64
        L.nlo_visibility_samples.argtypes = [fp, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32, C.c_int,
                                             fp, fp]
        L.nlo_throughput_samples.restype = C.c_int
        L.nlo_throughput_samples.argtypes = [fp, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64,
                                             C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, fp, fp, fp,
                                             u32p]
        L.nlo_downsample.argtypes = [u32p, C.c_uint32, u32p]
        L.nlo_march.argtypes = [u32p, C.c_uint32, fp, fp, C.c_uint32, fp, fp, C.c_void_p, C.c_void_p, u32p, fp, fp]
        L.nlo_optimal_thresholds.argtypes = [C.c_uint32, C.c_uint32, fp, fp, C.c_uint32, C.c_int, C.c_float, fp, fp]
        L.nlo_omp_max_threads.restype = C.c_int
        L.nlo_omp_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def mesh_bbox(tris: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    tris = np.ascontiguousarray(tris, dtype=np.float32)
    mn = np.zeros(3, np.float32)
    mx = np.zeros(3, np.float32)
    lib().nlo_mesh_bbox(_fptr(tris), tris.size // 9, _fptr(mn), _fptr(mx))
    return mn, mx


def voxinfo(bbox_min, bbox_max, res: int, pad: bool, n_tri: int) -> VoxInfo:
    mn = np.ascontiguousarray(bbox_min, dtype=np.float32)
    mx = np.ascontiguousarray(bbox_max, dtype=np.float32)
    info = VoxInfo()
    lib().nlo_voxinfo_from_bbox(_fptr(mn), _fptr(mx), res, 1e-4 if pad else 0.0, n_tri, C.byref(info))
    return info


def table_words(res: int) -> int:
    return (res ** 3 + 31) // 32


def voxelize(tris: np.ndarray, res: int, solid: bool = False, pad: bool = False, morton: bool = False,
             info: VoxInfo | None = None, threads: int | None = None) -> tuple[np.ndarray, VoxInfo]:
    """Bit table (uint32[ceil(res^3/32)]) of the reference CPU voxelizer."""
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    if info is None:
        mn, mx = mesh_bbox(tris)
        info = voxinfo(mn, mx, res, pad, tris.shape[0])
    table = np.zeros(table_words(res), np.uint32)
    L = lib()
    prev = L.nlo_omp_max_threads()
    if threads is not None:
        L.nlo_omp_set_threads(int(threads))
    try:
        fn = L.nlo_voxelize_solid if solid else L.nlo_voxelize_surface
        fn(C.byref(info), _fptr(tris), table.ctypes.data_as(C.POINTER(C.c_uint32)), int(morton))
    finally:
        if threads is not None:  # the thread count is process-wide: put it back
            L.nlo_omp_set_threads(int(prev))
    return table, info


def unpack_bits(table: np.ndarray, res: int) -> np.ndarray:
    """Occupancy as bool[z][y][x] from the MSB-first bit table."""
    bits = np.unpackbits(table.astype(">u4").view(np.uint8))[: res ** 3]
    return bits.reshape(res, res, res).astype(bool)


def visibility_samples(tris: np.ndarray, res: int, seed: int, first: int, n: int,
                       brute_force: bool = False) -> tuple[np.ndarray, np.ndarray, int]:
    """(inputs [n, 6], labels [n, 1], occupied voxel count) of the visibility
    sample stream; brute_force tests every triangle instead of the voxel's list."""
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    inputs = np.zeros((n, 6), np.float32)
    labels = np.zeros((n, 1), np.float32)
    n_occ = lib().nlo_visibility_samples(_fptr(tris), tris.shape[0], res, int(seed) & (2 ** 64 - 1), first, n,
                                         int(brute_force), _fptr(inputs), _fptr(labels))
    return inputs, labels, int(n_occ)


def throughput_samples(tris: np.ndarray, res: int, n_lods: int, seed: int, first: int, n: int, spp: int = 64,
                       max_depth: int = 3, albedo_seed: int = 1, brute_force: bool = False):
    """(inputs [n, 10], targets [n, 3], valid [n], occupied voxels per level) of
    the throughput sample stream."""
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    inputs = np.zeros((n, 10), np.float32)
    targets = np.zeros((n, 3), np.float32)
    valid = np.zeros(n, np.float32)
    n_occ = np.zeros(8, np.uint32)
    ok = lib().nlo_throughput_samples(_fptr(tris), tris.shape[0], res, n_lods, int(seed) & (2 ** 64 - 1), first, n,
                                      spp, max_depth, int(albedo_seed) & (2 ** 64 - 1), int(brute_force),
                                      _fptr(inputs), _fptr(targets), _fptr(valid),
                                      n_occ.ctypes.data_as(C.POINTER(C.c_uint32)))
    if not ok:
        raise ValueError("empty level or bad n_lods")
    return inputs, targets, valid, [int(v) for v in n_occ[:n_lods]]


def downsample(table: np.ndarray, res: int) -> np.ndarray:
    """Bit table of the (res/2)^3 level: OR of 2x2x2 blocks."""
    child = np.ascontiguousarray(table, dtype=np.uint32)
    parent = np.zeros(table_words(res // 2), np.uint32)
    u32 = C.POINTER(C.c_uint32)
    lib().nlo_downsample(child.ctypes.data_as(u32), res, parent.ctypes.data_as(u32))
    return parent


def march(table: np.ndarray, res: int, bmin, unit, origins: np.ndarray, dirs: np.ndarray, t_min=None, prev=None):
    """(voxel uint32 [n] with 0xffffffff for none, t_near [n], t_far [n])."""
    bits = np.ascontiguousarray(table, dtype=np.uint32)
    o = np.ascontiguousarray(origins, np.float32)
    d = np.ascontiguousarray(dirs, np.float32)
    n = o.shape[0]
    bm = np.ascontiguousarray(bmin, np.float32)
    un = np.ascontiguousarray(unit, np.float32)
    vox = np.zeros(n, np.uint32)
    tn = np.zeros(n, np.float32)
    tf = np.zeros(n, np.float32)
    tm = None if t_min is None else np.ascontiguousarray(t_min, np.float32)
    pv = None if prev is None else np.ascontiguousarray(prev, np.uint32)
    u32 = C.POINTER(C.c_uint32)
    lib().nlo_march(bits.ctypes.data_as(u32), res, _fptr(bm), _fptr(un), n, _fptr(o), _fptr(d),
                    None if tm is None else tm.ctypes.data, None if pv is None else pv.ctypes.data,
                    vox.ctypes.data_as(u32), _fptr(tn), _fptr(tf))
    return vox, tn, tf


def morton3(x, y, z) -> np.ndarray:
    """encode_morton3 (rptr/cuda/neural-lod/neural_lod_learning.h:519-532): the
    low 10 bits of x, y, z interleaved, x in bit 0, y in bit 1, z in bit 2."""
    def part(v):
        v = np.asarray(v, np.uint64) & np.uint64(0x3FF)
        v = (v ^ (v << np.uint64(16))) & np.uint64(0xFF0000FF)
        v = (v ^ (v << np.uint64(8))) & np.uint64(0x0300F00F)
        v = (v ^ (v << np.uint64(4))) & np.uint64(0x030C30C3)
        v = (v ^ (v << np.uint64(2))) & np.uint64(0x09249249)
        return v
    return ((part(z) << np.uint64(2)) + (part(y) << np.uint64(1)) + part(x)).astype(np.uint32)


def active_voxels(table: np.ndarray, res: int) -> np.ndarray:
    """Linear indices x + y R + z R^2 of the occupied voxels of a bit table."""
    bits = unpack_bits(np.ascontiguousarray(table, np.uint32), res).reshape(-1)
    return np.flatnonzero(bits)


def voxel_thresholds(table: np.ndarray, res: int, thresholds: np.ndarray, lo: float, hi: float) -> np.ndarray:
    """Dense float32 [R^3] thresholds as update_grid_with_thresholds leaves the
    level (neural_lod_learning.cu:50-121): the j-th occupied voxel in
    increasing Morton order (the sorted morton_sparse_voxel_idx_mappings,
    neural_lod_learning_init.cu:288-295) takes glm::clamp(t[j], lo, hi) =
    min(max(t[j], lo), hi); every other voxel keeps 0.5."""
    lin = active_voxels(table, res)
    if lin.size != len(thresholds):
        raise ValueError("one threshold per occupied voxel")
    x, y, z = lin % res, (lin // res) % res, lin // (res * res)
    order = np.argsort(morton3(x, y, z), kind="stable")
    t = np.asarray(thresholds, np.float32)
    dense = np.full(res ** 3, np.float32(0.5), np.float32)
    dense[lin[order]] = np.minimum(np.maximum(t, np.float32(lo)), np.float32(hi))
    return dense


def optimal_thresholds(labels: np.ndarray, probas: np.ndarray, n_voxels: int, n_thresholds: int = 101,
                       metric: int = 2, beta: float = 1.0):
    """(thresholds [V], scores [V][T]) of the per-voxel threshold optimisation."""
    lab = np.ascontiguousarray(labels, np.float32).ravel()
    pro = np.ascontiguousarray(probas, np.float32).ravel()
    S = lab.size // n_voxels
    thr = np.zeros(n_voxels, np.float32)
    sc = np.zeros((n_voxels, n_thresholds), np.float32)
    lib().nlo_optimal_thresholds(n_voxels, S, _fptr(lab), _fptr(pro), n_thresholds, metric, beta, _fptr(thr),
                                 _fptr(sc))
    return thr, sc
