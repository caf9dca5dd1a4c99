"""GPU per-voxel threshold optimisation (libnlod nlod_optimal_thresholds)
against the CPU restatement: identical scores and thresholds for every
metric."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _data(V, S, seed):
    rng = np.random.default_rng(seed)
    probas = rng.random((V, S)).astype(np.float32)
    probas[rng.random((V, S)) < 0.05] = np.float32(0.5)  # exactly on a threshold
    labels = (rng.random((V, S)) < probas * 0.7 + 0.15).astype(np.float32)
    labels[:3] = 0.0  # voxels without positives
    labels[3:5] = 1.0  # and without negatives
    return labels, probas


@pytest.mark.parametrize("T", [101, 100, 2])
def test_all_metrics_match_oracle(oracle, cuda, T):
    from neural_lod_amd.thresholds import optimal_thresholds
    V, S = 300, 96
    labels, probas = _data(V, S, T)
    lt, pt = torch.from_numpy(labels).cuda(), torch.from_numpy(probas).cuda()
    for metric in range(12):
        thr, sc = optimal_thresholds(lt, pt, V, T, metric=metric, beta=2.0, return_scores=True)
        ethr, esc = oracle.optimal_thresholds(labels, probas, V, T, metric=metric, beta=2.0)
        assert np.array_equal(sc.cpu().numpy(), esc), f"T={T} metric {metric}: scores differ"
        assert np.array_equal(thr.cpu().numpy(), ethr), f"T={T} metric {metric}: thresholds differ"


def test_large_voxels(oracle, cuda):
    from neural_lod_amd.thresholds import optimal_thresholds
    V, S = 7, 5000
    labels, probas = _data(V, S, 9)
    thr = optimal_thresholds(torch.from_numpy(labels).cuda(), torch.from_numpy(probas).cuda(), V)
    ethr, _ = oracle.optimal_thresholds(labels, probas, V, 101, metric=2)
    assert np.array_equal(thr.cpu().numpy(), ethr)


def test_invalid_arguments(cuda):
    from neural_lod_amd._lib import NlodError
    from neural_lod_amd.thresholds import optimal_thresholds
    x = torch.zeros(10, device="cuda")
    with pytest.raises(NlodError):
        optimal_thresholds(x, x, 2, n_thresholds=1)
    with pytest.raises(NlodError):
        optimal_thresholds(x, x, 2, metric=12)
