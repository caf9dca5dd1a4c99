"""Per-voxel visibility thresholds (libnlod ``nlod_optimal_thresholds``): the
threshold t / (T - 1) maximising a classification score of the visibility
network's outputs against ray-traced labels, voxel by voxel
(rptr/cuda/neural-lod/neural_lod_learning.cu:122-266)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

METRICS = {"precision": 0, "recall": 1, "fscore": 2, "fscore_macro": 3, "fscore_weighted": 4,
           "optimal_fscore_weighted": 5, "mcc": 6, "cohen_kappa": 7, "accuracy": 8, "balanced_accuracy": 9,
           "fowlkes_mallows": 10, "youden_j": 11}


def optimal_thresholds(labels: torch.Tensor, probas: torch.Tensor, n_voxels: int, n_thresholds: int = 101,
                       metric: str | int = "fscore", beta: float = 1.0, return_scores: bool = False):
    """labels / probas: float32 CUDA tensors of n_voxels x samples (voxel-major).
    Returns thresholds [n_voxels] (and scores [n_voxels, n_thresholds])."""
    m = METRICS[metric] if isinstance(metric, str) else int(metric)
    for t, name in ((labels, "labels"), (probas, "probas")):
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
            raise TypeError(f"{name} must be a contiguous float32 CUDA tensor")
    if labels.numel() != probas.numel() or n_voxels <= 0 or labels.numel() % n_voxels:
        raise ValueError("labels and probas must hold n_voxels x samples values")
    thr = torch.empty(n_voxels, dtype=torch.float32, device=labels.device)
    sc = torch.empty((n_voxels, n_thresholds), dtype=torch.float32, device=labels.device) if return_scores else None
    stream = C.c_void_p(torch.cuda.current_stream(labels.device).cuda_stream)
    _lib.call("nlod_optimal_thresholds", stream, n_voxels, labels.numel() // n_voxels, C.c_void_p(labels.data_ptr()),
              C.c_void_p(probas.data_ptr()), n_thresholds, m, float(beta), C.c_void_p(thr.data_ptr()),
              C.c_void_p(sc.data_ptr()) if sc is not None else None)
    return (thr, sc) if return_scores else thr


def threshold_path(weights_directory: str, weights_filename_prefix: str, lod: int) -> str:
    """<weights dir>../optimal_thresholds/<prefix>_lod_<l>_optimal_thresholds.bin
    (get_optimal_threshold_path, rptr/cuda/neural-lod/neural_lod_learning.h:332-334)."""
    return f"{weights_directory}../optimal_thresholds/{weights_filename_prefix}_lod_{int(lod)}_optimal_thresholds.bin"


def save_thresholds(path, thresholds) -> None:
    """Raw little-endian float32 values, nothing else (write_vector_to_file,
    neural_lod_learning.h:464-480)."""
    a = thresholds.detach().cpu().numpy() if isinstance(thresholds, torch.Tensor) else np.asarray(thresholds)
    np.ascontiguousarray(a, dtype="<f4").tofile(path)


def load_thresholds(path, n: int) -> np.ndarray:
    """The first n float32 values of a threshold file (read_vector_from_file,
    neural_lod_learning.h:482-496, which reads n = the level's occupied voxel
    count).  A file holding fewer than n values is an error here (the
    reference leaves the missing values zero)."""
    a = np.fromfile(path, dtype="<f4", count=int(n))
    if a.size != int(n):
        raise ValueError(f"{path}: {a.size} thresholds, expected {int(n)}")
    return a.astype(np.float32)

